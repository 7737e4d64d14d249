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
