// bvh_common.h -- pieces shared by the BVH builders (bvh_build.hip: LBVH,
// bvh_sah.hip: top-down binned SAH): ordered-float atomics, padded triangle
// boxes, the scene-bounds reduction and the TriRecord emission.
//
// Boxes are padded outwards (relative 2^-20 of the coordinate plus 2^-22 of
// the scene extent) so that box tests never cull a triangle that the exact
// Moller-Trumbore test accepts.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "internal.h"
#include "pt_math.h"

namespace mipt {
namespace bvhc {
// Internal linkage: each builder's translation unit has its own copy.
namespace {

constexpr float SAH_C_NODE = 1.0f;
constexpr float SAH_C_TRI = 1.0f;
constexpr float PAD_REL = 9.5367431640625e-07f;  // 2^-20
constexpr float PAD_ABS = 2.384185791015625e-07f; // 2^-22 of the scene extent

// float <-> uint32 whose unsigned order is the float order (atomicMin/Max on
// floats through integer atomics).
__host__ __device__ __forceinline__ uint32_t f2o(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ __forceinline__ float o2f(uint32_t u)
{
    uint32_t v = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    memcpy(&f, &v, 4);
    return f;
}

__device__ __forceinline__ float wave_min(float v)
{
    for (int off = 32; off > 0; off >>= 1)
        v = fminf(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v)
{
    for (int off = 32; off > 0; off >>= 1)
        v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

__host__ __device__ __forceinline__ float half_area(float4 lo, float4 hi)
{
    float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dy * dz + dz * dx;
}

__host__ __device__ __forceinline__ float4 f4min(float4 a, float4 b)
{
    return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.0f);
}
__host__ __device__ __forceinline__ float4 f4max(float4 a, float4 b)
{
    return make_float4(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z), 0.0f);
}

// The padded box of a triangle box [a, b].
__host__ __device__ __forceinline__ void pad_box(float4 a, float4 b, float abs_pad, float4& lo, float4& hi)
{
    lo = make_float4(a.x - (fabsf(a.x) * PAD_REL + abs_pad), a.y - (fabsf(a.y) * PAD_REL + abs_pad),
                     a.z - (fabsf(a.z) * PAD_REL + abs_pad), 0.0f);
    hi = make_float4(b.x + (fabsf(b.x) * PAD_REL + abs_pad), b.y + (fabsf(b.y) * PAD_REL + abs_pad),
                     b.z + (fabsf(b.z) * PAD_REL + abs_pad), 0.0f);
}

constexpr int BOUNDS_BLOCK = 256;
constexpr int BOUNDS_GRID_MAX = 1024;

// Triangle boxes plus per-block partial bounds (12 floats: scene lo/hi,
// centroid lo/hi). A grid-stride loop keeps the grid at most
// BOUNDS_GRID_MAX blocks; k_bounds_final folds the partials without atomics.
__global__ __launch_bounds__(BOUNDS_BLOCK) void k_tri_boxes(const float* __restrict__ verts,
                                                            const int32_t* __restrict__ idx, int32_t n,
                                                            float4* __restrict__ lo_out, float4* __restrict__ hi_out,
                                                            float* __restrict__ partial)
{
    float v[12];
    for (int k = 0; k < 12; k++) v[k] = (k % 6) < 3 ? INFINITY : -INFINITY;
    for (int32_t i = blockIdx.x * BOUNDS_BLOCK + threadIdx.x; i < n; i += gridDim.x * BOUNDS_BLOCK)
    {
        pt::f3 p[3];
        for (int k = 0; k < 3; k++)
        {
            int32_t vi = idx[3 * (int64_t)i + k];
            p[k] = pt::mk(verts[3 * (int64_t)vi], verts[3 * (int64_t)vi + 1], verts[3 * (int64_t)vi + 2]);
        }
        pt::f3 lo = pt::fmin3(p[0], pt::fmin3(p[1], p[2]));
        pt::f3 hi = pt::fmax3(p[0], pt::fmax3(p[1], p[2]));
        lo_out[i] = make_float4(lo.x, lo.y, lo.z, 0.0f);
        hi_out[i] = make_float4(hi.x, hi.y, hi.z, 0.0f);
        float cx = 0.5f * (lo.x + hi.x), cy = 0.5f * (lo.y + hi.y), cz = 0.5f * (lo.z + hi.z);
        float t[12] = {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, cx, cy, cz, cx, cy, cz};
        for (int k = 0; k < 12; k++) v[k] = (k % 6) < 3 ? fminf(v[k], t[k]) : fmaxf(v[k], t[k]);
    }
    __shared__ float red[BOUNDS_BLOCK / 64][12];
    for (int k = 0; k < 12; k++)
        v[k] = ((k % 6) < 3) ? wave_min(v[k]) : wave_max(v[k]);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 12; k++) red[threadIdx.x / 64][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 12)
    {
        const int k = threadIdx.x;
        float r = red[0][k];
        for (int w = 1; w < BOUNDS_BLOCK / 64; w++) r = (k % 6) < 3 ? fminf(r, red[w][k]) : fmaxf(r, red[w][k]);
        partial[12 * blockIdx.x + k] = r;
    }
}

// Folds the partials of `blocks` blocks into bounds (as ordered uint32,
// [0..2] scene lo, [3..5] scene hi, [6..8] centroid lo, [9..11] centroid hi)
// and the absolute pad (2^-22 of the largest scene extent).
__global__ void k_bounds_final(const float* __restrict__ partial, int32_t blocks, uint32_t* __restrict__ bounds,
                               float* __restrict__ pad)
{
    __shared__ float res[12];
    if (threadIdx.x < 12)
    {
        const int k = threadIdx.x;
        float r = (k % 6) < 3 ? INFINITY : -INFINITY;
        for (int32_t b = 0; b < blocks; b++)
            r = (k % 6) < 3 ? fminf(r, partial[12 * b + k]) : fmaxf(r, partial[12 * b + k]);
        res[k] = r;
        bounds[k] = f2o(r);
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        float ext = 0.0f;
        for (int k = 0; k < 3; k++) ext = fmaxf(ext, res[3 + k] - res[k]);
        pad[0] = ext * PAD_ABS;
    }
}

// 48-byte triangle records in leaf order: record k holds primitive vals[k].
__global__ void k_emit_tris(const float* __restrict__ verts, const int32_t* __restrict__ idx, int32_t n,
                            const uint32_t* __restrict__ vals, TriRecord* __restrict__ tris)
{
    int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    int32_t prim = (int32_t)vals[k];
    pt::f3 p[3];
    for (int j = 0; j < 3; j++)
    {
        int32_t vi = idx[3 * (int64_t)prim + j];
        p[j] = pt::mk(verts[3 * (int64_t)vi], verts[3 * (int64_t)vi + 1], verts[3 * (int64_t)vi + 2]);
    }
    pt::f3 e1 = p[1] - p[0];
    pt::f3 e2 = p[2] - p[0];
    TriRecord r;
    r.v0 = make_float4(p[0].x, p[0].y, p[0].z, __int_as_float(prim));
    r.e1 = make_float4(e1.x, e1.y, e1.z, 0.0f);
    r.e2 = make_float4(e2.x, e2.y, e2.z, 0.0f);
    tris[k] = r;
}

inline unsigned blocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

inline unsigned bounds_grid(int32_t n)
{
    const unsigned g = blocks(n, BOUNDS_BLOCK);
    return g < (unsigned)BOUNDS_GRID_MAX ? (g ? g : 1u) : (unsigned)BOUNDS_GRID_MAX;
}

} // namespace
} // namespace bvhc

// Host side, after a builder wrote `node_capacity` BvhNodes (root = node 0)
// and the TriRecords: validates the reachable tree (each interior node
// reached once, each triangle in exactly one leaf), computes its statistics
// and SAH cost, and lays the reachable nodes out depth first. Defined in
// bvh_build.hip.
void finalize_bvh(Ctx& ctx, int32_t node_capacity);

// The one-triangle tree: a root whose first child is the leaf and whose
// second child has an empty box. Defined in bvh_build.hip.
void write_single_triangle_root(Ctx& ctx, const float4* d_lo, const float4* d_hi);

} // namespace mipt
