/*
 * fmath.h -- sin, cos and pow of the oracle (test infrastructure).
 *
 * The reference calls the platform's cosf/sinf/powf, whose last-place rounding
 * differs between a GPU math library and glibc. The oracle and the GPU kernels
 * therefore both evaluate these functions by one fixed sequence of binary32
 * operations, written out independently on each side: Cody & Waite reduction
 * by pi/2 in three parts, then Taylor polynomials (sin to r^9, cos to r^10) on
 * |r| <= pi/4; log2 through ln m = 2 atanh((m-1)/(m+1)); exp2 through the
 * exponential series to g^7; integral powers up to 64 by square and multiply.
 * tests/test_oracle_fmath.py checks each against the correctly rounded value.
 * Compiled with -ffp-contract=off like the rest of the oracle.
 */
#ifndef MIPT_ORACLE_FMATH_H
#define MIPT_ORACLE_FMATH_H

#include <math.h>

static inline void orc_sincos(float x, float* s_out, float* c_out)
{
    float k = rintf(x * 0x1.45f306p-1f);            /* 2 / pi */
    float r = x - k * 0x1.92p+0f;                   /* 1.5703125 */
    r = r - k * 0x1.fb6p-12f;
    r = r - k * -0x1.777a5cp-25f;
    float r2 = r * r;
    float sp = r + r * r2 * (-0x1.555556p-3f + r2 * (0x1.111112p-7f + r2 * (-0x1.a01a02p-13f + r2 * 0x1.71de3ap-19f)));
    float cp = 1.0f + r2 * (-0.5f + r2 * (0x1.555556p-5f + r2 * (-0x1.6c16c2p-10f + r2 * (0x1.a01a02p-16f + r2 * -0x1.27e4fcp-22f))));
    int q = (int)k & 3;
    *s_out = q == 0 ? sp : q == 1 ? cp : q == 2 ? -sp : -cp;
    *c_out = q == 0 ? cp : q == 1 ? -sp : q == 2 ? -cp : sp;
}

static inline float orc_sinf(float x)
{
    float s, c;
    orc_sincos(x, &s, &c);
    return s;
}

static inline float orc_cosf(float x)
{
    float s, c;
    orc_sincos(x, &s, &c);
    return c;
}

static inline float orc_log2f(float x)
{
    if (x == 0.0f)
        return -INFINITY;
    int e;
    float m = frexpf(x, &e);
    if (m < 0x1.6a09e6p-1f) /* sqrt(1/2) */
    {
        m = m * 2.0f;
        e = e - 1;
    }
    float t = (m - 1.0f) / (m + 1.0f);
    float t2 = t * t;
    float ln_m = t * (2.0f + t2 * (0x1.555556p-1f + t2 * (0x1.99999ap-2f + t2 * (0x1.24924ap-2f + t2 * 0x1.c71c72p-3f))));
    return (float)e + ln_m * 0x1.715476p+0f; /* 1 / ln 2 */
}

static inline float orc_exp2f(float v)
{
    if (v < -150.0f)
        return 0.0f;
    if (v > 128.0f)
        return INFINITY;
    float n = rintf(v);
    float g = (v - n) * 0x1.62e43p-1f; /* ln 2 */
    float p = 1.0f + g * (1.0f + g * (0.5f + g * (0x1.555556p-3f + g * (0x1.555556p-5f + g * (0x1.111112p-7f + g * (0x1.6c16c2p-10f + g * 0x1.a01a02p-13f))))));
    return ldexpf(p, (int)n);
}

static inline float orc_powf(float x, float y)
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
    return orc_exp2f(y * orc_log2f(x));
}

/* Natural logarithm and exponential through the base-2 sequences. */
static inline float orc_logf(float x) { return orc_log2f(x) * 0x1.62e43p-1f; }
static inline float orc_expf(float x) { return orc_exp2f(x * 0x1.715476p+0f); }

/* atan2 with IEEE zero and sign handling: the quotient of the smaller by the
 * larger magnitude is reduced to |t| <= tan(pi/8) with
 * atan(a) = pi/4 + atan((a - 1) / (a + 1)), then the odd Taylor series of atan
 * to t^15. */
static inline float orc_atan2f(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    float r;
    if (ax == 0.0f && ay == 0.0f)
        r = signbit(x) ? 0x1.921fb6p+1f : 0.0f;
    else
    {
        const int swap = ay > ax;
        const float a = swap ? ax / ay : ay / ax;
        float base = 0.0f, t = a;
        if (a > 0x1.a8279ap-2f)
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

/* asin and acos through atan2 of the legs x and sqrt((1 - x)(1 + x)), the
 * argument clamped to [-1, 1]. */
static inline float orc_asinf(float y)
{
    y = fminf(fmaxf(y, -1.0f), 1.0f);
    return orc_atan2f(y, sqrtf((1.0f - y) * (1.0f + y)));
}

static inline float orc_acosf(float x)
{
    x = fminf(fmaxf(x, -1.0f), 1.0f);
    return orc_atan2f(sqrtf((1.0f - x) * (1.0f + x)), x);
}

#endif
