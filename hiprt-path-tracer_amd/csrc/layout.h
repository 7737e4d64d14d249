// layout.h -- HBM layout of the acceleration structure.
//
// BvhNode: one 64-byte record per interior node of a binary BVH = one cache
// line. It carries the boxes of BOTH children, so a traversal step is one
// 64-byte fetch that decides which children to visit:
//   a = (c0.lo.x, c0.lo.y, c0.lo.z, c0.hi.x)
//   b = (c0.hi.y, c0.hi.z, c1.lo.x, c1.lo.y)
//   c = (c1.lo.z, c1.hi.x, c1.hi.y, c1.hi.z)
//   d = (child0, child1, 0, 0)
// A child >= 0 is an interior node index; a child < 0 is a leaf
// ~((first << LEAF_COUNT_BITS) | (count - 1)) over TriRecords [first, first+count).
//
// TriRecord: 48 bytes per triangle in leaf order: (a, prim), (b - a, 0),
// (c - a, 0). The edges are the exact binary32 differences the reference
// computes per test (src/Renderer/Triangle.h:26-27).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mipt {

constexpr int LEAF_COUNT_BITS = 4;
constexpr int MAX_LEAF_SIZE = 1 << LEAF_COUNT_BITS; // 16

struct alignas(64) BvhNode {
    float4 a, b, c;
    int4 d;
};
static_assert(sizeof(BvhNode) == 64, "BvhNode must be one 64-byte line");

struct alignas(16) TriRecord {
    float4 v0; // w = primitive index (int bits)
    float4 e1;
    float4 e2;
};
static_assert(sizeof(TriRecord) == 48, "TriRecord must be 48 bytes");

__host__ __device__ inline int32_t make_leaf(int32_t first, int32_t count)
{
    return ~((first << LEAF_COUNT_BITS) | (count - 1));
}
__host__ __device__ inline int32_t leaf_first(int32_t child) { return (~child) >> LEAF_COUNT_BITS; }
__host__ __device__ inline int32_t leaf_count(int32_t child) { return ((~child) & (MAX_LEAF_SIZE - 1)) + 1; }

} // namespace mipt
