// pt_math.h -- float3 arithmetic, RNG and the ray/triangle test shared by all
// kernels of libmipt. Everything here is compiled with -ffp-contract=off: each
// expression is evaluated as written, one IEEE binary32 operation at a time,
// with correctly rounded division and square root (HIP's defaults).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PT_HD __host__ __device__ __forceinline__

namespace pt {

struct f3 {
    float x, y, z;
};

PT_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
PT_HD f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
PT_HD f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
PT_HD f3 operator*(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
PT_HD f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
// dot = (x*x + y*y) + z*z, left to right.
PT_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PT_HD f3 cross(f3 a, f3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// normalize divides by the correctly rounded length (no reciprocal).
PT_HD f3 normalize(f3 a)
{
    float l = sqrtf(dot(a, a));
    return mk(a.x / l, a.y / l, a.z / l);
}
PT_HD f3 fmin3(f3 a, f3 b) { return mk(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
PT_HD f3 fmax3(f3 a, f3 b) { return mk(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }

// ---------------------------------------------------------------------------
// sin, cos and pow as a fixed sequence of binary32 operations. The device
// library and a host libm round sinf/cosf/powf differently in the last
// place; with these, the kernels and any host restatement of the same steps
// produce identical bits. Accuracy: a few ulp for |x| < 6000 (sin, cos) and
// for x >= 0 (pow); the renderer passes angles in [0, 2 pi] and x in [0, 1].
// ---------------------------------------------------------------------------

// Cody & Waite reduction x = k pi/2 + r with pi/2 split in three parts
// (1.5703125 and 0x1.fb6p-12 have 8 and 12 significant bits, so k times
// either is exact for |k| < 4096), then the Taylor polynomials of sin and
// cos on |r| <= pi/4 (truncation error < 2e-9).
PT_HD void sincos_fixed(float x, float& s_out, float& c_out)
{
    const float k = rintf(x * 0x1.45f306p-1f);
    float r = x - k * 0x1.92p+0f;
    r = r - k * 0x1.fb6p-12f;
    r = r - k * -0x1.777a5cp-25f;
    const float r2 = r * r;
    const float sp = r + r * r2 * (-0x1.555556p-3f + r2 * (0x1.111112p-7f + r2 * (-0x1.a01a02p-13f + r2 * 0x1.71de3ap-19f)));
    const float cp = 1.0f + r2 * (-0.5f + r2 * (0x1.555556p-5f + r2 * (-0x1.6c16c2p-10f + r2 * (0x1.a01a02p-16f + r2 * -0x1.27e4fcp-22f))));
    const int q = (int)k & 3;
    s_out = q == 0 ? sp : q == 1 ? cp : q == 2 ? -sp : -cp;
    c_out = q == 0 ? cp : q == 1 ? -sp : q == 2 ? -cp : sp;
}
PT_HD float sin_fixed(float x)
{
    float s, c;
    sincos_fixed(x, s, c);
    return s;
}
PT_HD float cos_fixed(float x)
{
    float s, c;
    sincos_fixed(x, s, c);
    return c;
}

// log2 of a positive normal x: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(t), t = (m - 1) / (m + 1), |t| < 0.172, by its series to t^9;
// log2(0) = -inf.
PT_HD float log2_fixed(float x)
{
    if (x == 0.0f)
        return -INFINITY;
    int e;
    float m = frexpf(x, &e);
    if (m < 0x1.6a09e6p-1f)
    {
        m = m * 2.0f;
        e = e - 1;
    }
    const float t = (m - 1.0f) / (m + 1.0f);
    const float t2 = t * t;
    const float ln_m = t * (2.0f + t2 * (0x1.555556p-1f + t2 * (0x1.99999ap-2f + t2 * (0x1.24924ap-2f + t2 * 0x1.c71c72p-3f))));
    return (float)e + ln_m * 0x1.715476p+0f;
}

// 2^v = 2^n e^(f ln 2), n = rint(v), |f| <= 1/2, by the Taylor series of exp to g^7.
PT_HD float exp2_fixed(float v)
{
    if (v < -150.0f)
        return 0.0f;
    if (v > 128.0f)
        return INFINITY;
    const float n = rintf(v);
    const float g = (v - n) * 0x1.62e43p-1f;
    const float p = 1.0f + g * (1.0f + g * (0.5f + g * (0x1.555556p-3f + g * (0x1.555556p-5f + g * (0x1.111112p-7f + g * (0x1.6c16c2p-10f + g * 0x1.a01a02p-13f))))));
    return ldexpf(p, (int)n);
}

// x^y for x >= 0. Integral exponents up to 64 in magnitude by square and
// multiply (the F82 Fresnel falloff exponent is 5 by default).
PT_HD float pow_fixed(float x, float y)
{
    if (y == rintf(y) && fabsf(y) <= 64.0f)
    {
        int n = (int)fabsf(y);
        float r = 1.0f, b = x;
        while (n)
        {
            if (n & 1)
                r = r * b;
            b = b * b;
            n >>= 1;
        }
        return y < 0.0f ? 1.0f / r : r;
    }
    if (x == 0.0f)
        return y > 0.0f ? 0.0f : INFINITY;
    if (x == 1.0f)
        return 1.0f;
    return exp2_fixed(y * log2_fixed(x));
}

// Natural logarithm and exponential through the base-2 sequences.
PT_HD float log_fixed(float x) { return log2_fixed(x) * 0x1.62e43p-1f; }
PT_HD float exp_fixed(float x) { return exp2_fixed(x * 0x1.715476p+0f); }

// atan2 with IEEE zero and sign handling: the quotient of the smaller by the
// larger magnitude is reduced to |t| <= tan(pi/8) with
// atan(a) = pi/4 + atan((a - 1) / (a + 1)), then the odd Taylor series of
// atan to t^15 (truncation error below 2e-8).
PT_HD float atan2_fixed(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    float r;
    if (ax == 0.0f && ay == 0.0f)
        r = signbit(x) ? 0x1.921fb6p+1f : 0.0f;
    else
    {
        const bool swap = ay > ax;
        const float a = swap ? ax / ay : ay / ax;
        float base = 0.0f, t = a;
        if (a > 0x1.a8279ap-2f) // tan(pi/8)
        {
            t = (a - 1.0f) / (a + 1.0f);
            base = 0x1.921fb6p-1f;
        }
        const float t2 = t * t;
        const float p = t + t * (t2 * (-0x1.555556p-2f + t2 * (0x1.99999ap-3f + t2 * (-0x1.24924ap-3f + t2 * (0x1.c71c72p-4f + t2 * (-0x1.745d18p-4f + t2 * (0x1.3b13b2p-4f + t2 * -0x1.111112p-4f)))))));
        r = base + p;
        if (swap)
            r = 0x1.921fb6p+0f - r;
        if (signbit(x))
            r = 0x1.921fb6p+1f - r;
    }
    return signbit(y) ? -r : r;
}

// asin and acos as the angle of a right triangle with legs x and
// sqrt((1 - x)(1 + x)), through atan2_fixed; the argument is clamped to
// [-1, 1] first (a normalised direction can exceed 1 by an ulp).
PT_HD float asin_fixed(float y)
{
    y = fminf(fmaxf(y, -1.0f), 1.0f);
    return atan2_fixed(y, sqrtf((1.0f - y) * (1.0f + y)));
}
PT_HD float acos_fixed(float x)
{
    x = fminf(fmaxf(x, -1.0f), 1.0f);
    return atan2_fixed(sqrtf((1.0f - x) * (1.0f + x)), x);
}

// ---------------------------------------------------------------------------
// Shared-exponent RGB, R9G9B9E5 (RGBE9995Packed, src/HostDeviceCommon/
// Packing.h:441-485): channel q in [0, 511] in bits 0-8 (r), 9-17 (g), 18-26
// (b), exponent E in bits 27-31, value q 2^(E - 24). Channels are clamped to
// [0, 511/256 * 2^15]; the largest, floored at 2^-16, has binary exponent e,
// and E = e + 16, one more when that channel rounds up to 512 units; each
// channel is then q = round-half-even(c 2^(24 - E)) (a scale by a power of
// two, exact).
// ---------------------------------------------------------------------------
PT_HD uint32_t rgb9e5_pack(float r, float g, float b)
{
    const float hi = 0x1.ffp+15f;
    r = fminf(fmaxf(r, 0.0f), hi);
    g = fminf(fmaxf(g, 0.0f), hi);
    b = fminf(fmaxf(b, 0.0f), hi);
    const float m = fmaxf(fmaxf(0x1p-16f, r), fmaxf(g, b));
    uint32_t bits;
    __builtin_memcpy(&bits, &m, 4);
    int32_t E = (int32_t)(bits >> 23) - 127 + 16;
    // 0x7fc000: mantissa at or above 511.5 / 256 of the binade rounds to 512
    if ((bits & 0x7fffffu) >= 0x7fc000u)
        E += 1;
    const uint32_t scale_bits = (uint32_t)(24 - E + 127) << 23;
    float scale;
    __builtin_memcpy(&scale, &scale_bits, 4);
    const uint32_t qr = (uint32_t)rintf(r * scale), qg = (uint32_t)rintf(g * scale), qb = (uint32_t)rintf(b * scale);
    return ((uint32_t)E << 27) | (qb << 18) | (qg << 9) | qr;
}

PT_HD void rgb9e5_unpack(uint32_t p, float& r, float& g, float& b)
{
    const uint32_t scale_bits = (uint32_t)((int32_t)(p >> 27) - 24 + 127) << 23;
    float scale;
    __builtin_memcpy(&scale, &scale_bits, 4);
    r = (float)(p & 0x1ffu) * scale;
    g = (float)((p >> 9) & 0x1ffu) * scale;
    b = (float)((p >> 18) & 0x1ffu) * scale;
}

// ---------------------------------------------------------------------------
// RNG: the reference's per-pixel streams (src/Device/includes/Hash.h:11-19,
// src/HostDeviceCommon/Xorshift.h:18-66).
// ---------------------------------------------------------------------------
PT_HD uint32_t wang_hash(uint32_t s)
{
    s = (s ^ 61u) ^ (s >> 16);
    s *= 9u;
    s = s ^ (s >> 4);
    s *= 0x27d4eb2du;
    s = s ^ (s >> 15);
    return s;
}

struct Xorshift32 {
    uint32_t state;
    PT_HD uint32_t next()
    {
        uint32_t x = state;
        x ^= x << 13;
        x ^= x >> 17;
        x ^= x << 5;
        state = x;
        return x;
    }
    // float(x) / float(0xffffffff); the divisor rounds to 2^32 in binary32.
    PT_HD float operator()()
    {
        float a = (float)next() / 4294967296.0f;
        const float hi = 1.0f - 1.0e-7f;
        return a < hi ? a : hi;
    }
    PT_HD int32_t random_index(int32_t n)
    {
        int32_t r = (int32_t)((float)next() / 4294967296.0f * (float)n);
        return r < n - 1 ? r : n - 1;
    }
};

PT_HD uint32_t pixel_seed(uint32_t pix, uint32_t sample_number, uint32_t random_number, bool freeze)
{
    return freeze ? wang_hash(pix + 1u) : wang_hash((pix + 1u) * (sample_number + 1u) * random_number);
}

// ---------------------------------------------------------------------------
// Moller-Trumbore with the reference's constants (src/Renderer/Triangle.h:20-63),
// on pre-computed edges e1 = b - a, e2 = c - a (identical bits to computing
// them per test).
// ---------------------------------------------------------------------------
PT_HD bool mt_intersect(f3 a, f3 e1, f3 e2, f3 o, f3 d, float& t_out, float& u_out, float& v_out)
{
    const float EPS = 0.0000001f;
    f3 h = cross(d, e2);
    float det = dot(e1, h);
    if (det > -EPS && det < EPS)
        return false;
    float f = 1.0f / det;
    f3 s = o - a;
    float u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f)
        return false;
    f3 q = cross(s, e1);
    float v = f * dot(d, q);
    if (v < 0.0f || u + v > 1.0f)
        return false;
    float t = f * dot(e2, q);
    if (!(t > EPS))
        return false;
    t_out = t;
    u_out = u;
    v_out = v;
    return true;
}

} // namespace pt
