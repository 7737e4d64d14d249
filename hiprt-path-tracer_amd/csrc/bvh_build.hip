// bvh_build.hip -- GPU LBVH builder (replaces hiprtBuildGeometry,
// src/HIPRT-Orochi/HIPRTScene.h:87-150, whose implementation lives in the
// absent HIPRT library).
//
// Pipeline (all on the device, one stream):
//   1. per-triangle boxes + scene/centroid bounds (block partials, folded
//      without atomics)
//   2. 63-bit Morton code of each box centre (21 bits per axis)
//   3. radix sort (Morton, primitive) pairs
//   4. Karras 2012 hierarchy: node i of n-1 interior nodes from the key prefixes
//   5. bottom-up refit with arrival counters; each node also gets its SAH cost
//      and a collapse flag (subtree of <= MAX_LEAF_SIZE triangles becomes one
//      leaf when that is cheaper)
//   6. emit 64-byte BvhNodes and 48-byte TriRecords in Morton order.
// Boxes are padded (bvh_common.h). finalize_bvh (host) validates the tree,
// computes its statistics and lays it out depth first for both builders.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "bvh_common.h"

namespace mipt {
namespace {

using bvhc::f2o;
using bvhc::f4max;
using bvhc::f4min;
using bvhc::half_area;
using bvhc::o2f;
using bvhc::SAH_C_NODE;
using bvhc::SAH_C_TRI;

__device__ __forceinline__ uint64_t expand21(uint64_t v)
{
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__global__ void k_morton(const float4* __restrict__ lo, const float4* __restrict__ hi, int32_t n,
                         const uint32_t* __restrict__ bounds, uint64_t* __restrict__ keys,
                         uint32_t* __restrict__ vals)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float clo[3], cext[3];
    for (int k = 0; k < 3; k++)
    {
        clo[k] = o2f(bounds[6 + k]);
        float e = o2f(bounds[9 + k]) - clo[k];
        cext[k] = e > 0.0f ? e : 1.0f;
    }
    float4 a = lo[i], b = hi[i];
    float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
    uint64_t code = 0;
    for (int k = 0; k < 3; k++)
    {
        float f = (c[k] - clo[k]) / cext[k];
        f = fminf(fmaxf(f, 0.0f), 1.0f);
        uint64_t q = (uint64_t)fminf(f * 2097152.0f, 2097151.0f);
        code |= expand21(q) << (2 - k);
    }
    keys[i] = code;
    vals[i] = (uint32_t)i;
}

__device__ __forceinline__ int delta(const uint64_t* __restrict__ keys, int32_t n, int32_t i, int32_t j)
{
    if (j < 0 || j >= n)
        return -1;
    uint64_t x = keys[i] ^ keys[j];
    if (x == 0)
        return 64 + __clz((uint32_t)(i ^ j));
    return __clzll(x);
}

// Karras 2012, "Maximizing Parallelism in the Construction of BVHs, Octrees,
// and k-d Trees", Figure 4. Child refs: >= 0 interior node, < 0 ~leaf index.
__global__ void k_karras(const uint64_t* __restrict__ keys, int32_t n, int32_t* __restrict__ child,
                         int32_t* __restrict__ node_parent, int32_t* __restrict__ leaf_parent,
                         int32_t* __restrict__ range_lo, int32_t* __restrict__ range_hi)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1)
        return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int32_t lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin)
        lmax *= 2;
    int32_t l = 0;
    for (int32_t t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, i, i + (l + t) * d) > dmin)
            l += t;
    int32_t j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int32_t s = 0;
    int32_t t = l;
    do
    {
        t = (t + 1) / 2;
        if (delta(keys, n, i, i + (s + t) * d) > dnode)
            s += t;
    } while (t > 1);
    int32_t gamma = i + s * d + (d < 0 ? -1 : 0);
    int32_t first = i < j ? i : j;
    int32_t last = i < j ? j : i;
    int32_t left = (first == gamma) ? ~gamma : gamma;
    int32_t right = (last == gamma + 1) ? ~(gamma + 1) : gamma + 1;
    child[2 * i] = left;
    child[2 * i + 1] = right;
    if (left < 0) leaf_parent[~left] = i; else node_parent[left] = i;
    if (right < 0) leaf_parent[~right] = i; else node_parent[right] = i;
    range_lo[i] = first;
    range_hi[i] = last;
    if (i == 0)
        node_parent[0] = -1;
}

// Padded box of sorted leaf k.
__device__ __forceinline__ void leaf_box(const float4* __restrict__ plo, const float4* __restrict__ phi,
                                         const uint32_t* __restrict__ vals, int32_t k, float abs_pad,
                                         float4& lo, float4& hi)
{
    const uint32_t p = vals[k];
    bvhc::pad_box(plo[p], phi[p], abs_pad, lo, hi);
}

__global__ void k_refit(int32_t n, const float4* __restrict__ plo, const float4* __restrict__ phi,
                        const uint32_t* __restrict__ vals, const int32_t* __restrict__ child,
                        const int32_t* __restrict__ node_parent, const int32_t* __restrict__ leaf_parent,
                        const int32_t* __restrict__ range_lo, const int32_t* __restrict__ range_hi,
                        uint32_t* __restrict__ visits, float4* __restrict__ nlo, float4* __restrict__ nhi,
                        float* __restrict__ cost, uint8_t* __restrict__ collapse, const float* __restrict__ pad)
{
    int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    const float abs_pad = pad[0];
    int32_t p = leaf_parent[k];
    // A binary radix tree over distinct 96-bit (key, index) prefixes is at
    // most 96 levels deep; the bound only guards against a malformed tree.
    for (int level = 0; p >= 0 && p < n - 1 && level < 128; level++)
    {
        __threadfence();
        if (atomicAdd(&visits[p], 1u) == 0)
            return; // the other child is not done yet; its thread continues
        __threadfence();
        float4 lo[2], hi[2];
        float c[2];
        for (int s = 0; s < 2; s++)
        {
            int32_t ch = child[2 * p + s];
            if (ch < 0)
            {
                leaf_box(plo, phi, vals, ~ch, abs_pad, lo[s], hi[s]);
                c[s] = half_area(lo[s], hi[s]) * SAH_C_TRI;
            }
            else
            {
                lo[s] = nlo[ch];
                hi[s] = nhi[ch];
                c[s] = cost[ch];
            }
        }
        float4 blo = f4min(lo[0], lo[1]), bhi = f4max(hi[0], hi[1]);
        float a = half_area(blo, bhi);
        int32_t count = range_hi[p] - range_lo[p] + 1;
        float c_split = a * SAH_C_NODE + c[0] + c[1];
        float c_leaf = a * (float)count * SAH_C_TRI;
        bool as_leaf = count <= MAX_LEAF_SIZE && c_leaf <= c_split;
        nlo[p] = blo;
        nhi[p] = bhi;
        cost[p] = as_leaf ? c_leaf : c_split;
        collapse[p] = as_leaf ? 1 : 0;
        p = node_parent[p];
    }
}

__global__ void k_emit_nodes(int32_t n, const float4* __restrict__ plo, const float4* __restrict__ phi,
                             const uint32_t* __restrict__ vals, const int32_t* __restrict__ child,
                             const int32_t* __restrict__ range_lo, const int32_t* __restrict__ range_hi,
                             const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                             const uint8_t* __restrict__ collapse, const float* __restrict__ pad,
                             BvhNode* __restrict__ nodes)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1)
        return;
    float4 lo[2], hi[2];
    int32_t ref[2];
    for (int s = 0; s < 2; s++)
    {
        int32_t ch = child[2 * i + s];
        if (ch < 0)
        {
            leaf_box(plo, phi, vals, ~ch, pad[0], lo[s], hi[s]);
            ref[s] = make_leaf(~ch, 1);
        }
        else
        {
            lo[s] = nlo[ch];
            hi[s] = nhi[ch];
            ref[s] = collapse[ch] ? make_leaf(range_lo[ch], range_hi[ch] - range_lo[ch] + 1) : ch;
        }
    }
    BvhNode nd;
    nd.a = make_float4(lo[0].x, lo[0].y, lo[0].z, hi[0].x);
    nd.b = make_float4(hi[0].y, hi[0].z, lo[1].x, lo[1].y);
    nd.c = make_float4(lo[1].z, hi[1].x, hi[1].y, hi[1].z);
    nd.d = make_int4(ref[0], ref[1], 0, 0);
    nodes[i] = nd;
}

using bvhc::blocks;

} // namespace

void write_single_triangle_root(Ctx& ctx, const float4* d_lo, const float4* d_hi)
{
    float4 lo, hi;
    HIP_CHECK(hipMemcpy(&lo, d_lo, sizeof(float4), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(&hi, d_hi, sizeof(float4), hipMemcpyDeviceToHost));
    const float ext = std::max({hi.x - lo.x, hi.y - lo.y, hi.z - lo.z});
    float4 plo, phi;
    bvhc::pad_box(lo, hi, ext * bvhc::PAD_ABS, plo, phi);
    BvhNode nd;
    nd.a = make_float4(plo.x, plo.y, plo.z, phi.x);
    // Empty second child: every slab of an all-(+inf) box is at +inf or -inf,
    // so the box test always fails.
    nd.b = make_float4(phi.y, phi.z, INFINITY, INFINITY);
    nd.c = make_float4(INFINITY, INFINITY, INFINITY, INFINITY);
    nd.d = make_int4(make_leaf(0, 1), make_leaf(0, 1), 0, 0);
    HIP_CHECK(hipMemcpy(ctx.bvh.nodes.ptr, &nd, sizeof(nd), hipMemcpyHostToDevice));
}

void finalize_bvh(Ctx& ctx, int32_t ninner)
{
    Bvh& bvh = ctx.bvh;
    const int32_t n = bvh.triangle_count;
    bvh.root = 0;
    std::vector<BvhNode> hn(ninner);
    HIP_CHECK(hipMemcpy(hn.data(), bvh.nodes.ptr, sizeof(BvhNode) * ninner, hipMemcpyDeviceToHost));
    int32_t inner = 0, leaves = 0, maxd = 0, maxleaf = 0;
    double sah = 0.0, root_area = 0.0;
    struct Item { int32_t node; int32_t depth; };
    std::vector<Item> stack;
    std::vector<uint8_t> seen(ninner, 0);
    std::vector<uint8_t> tri_seen(n, 0);
    stack.push_back({0, 0});
    auto harea = [](float lx, float ly, float lz, float hx, float hy, float hz) {
        double dx = (double)hx - lx, dy = (double)hy - ly, dz = (double)hz - lz;
        if (dx < 0 || dy < 0 || dz < 0) return 0.0;
        return dx * dy + dy * dz + dz * dx;
    };
    {
        const BvhNode& r = hn[0];
        float lx = std::min(r.a.x, r.b.z), ly = std::min(r.a.y, r.b.w), lz = std::min(r.a.z, r.c.x);
        float hx = std::max(r.a.w, r.c.y), hy = std::max(r.b.x, r.c.z), hz = std::max(r.b.y, r.c.w);
        root_area = harea(lx, ly, lz, hx, hy, hz);
        if (n == 1) root_area = harea(r.a.x, r.a.y, r.a.z, r.a.w, r.b.x, r.b.y);
    }
    while (!stack.empty())
    {
        Item it = stack.back();
        stack.pop_back();
        if (it.node < 0 || it.node >= ninner || seen[it.node])
            fail(MIPT_ERR_HIP, "BVH validation failed: interior node reached twice or out of range");
        seen[it.node] = 1;
        inner++;
        const BvhNode& nd = hn[it.node];
        double areas[2] = {harea(nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y),
                           harea(nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w)};
        sah += SAH_C_NODE * (it.node == 0 ? root_area : 0.0);
        int32_t refs[2] = {nd.d.x, nd.d.y};
        int nchild = n == 1 ? 1 : 2;
        for (int c = 0; c < nchild; c++)
        {
            if (refs[c] < 0)
            {
                leaves++;
                int32_t cnt = leaf_count(refs[c]);
                int32_t first = leaf_first(refs[c]);
                if (first < 0 || first + cnt > n)
                    fail(MIPT_ERR_HIP, "BVH validation failed: leaf range out of bounds");
                for (int32_t q = first; q < first + cnt; q++)
                {
                    if (tri_seen[q] && n > 1)
                        fail(MIPT_ERR_HIP, "BVH validation failed: triangle in two leaves");
                    tri_seen[q] = 1;
                }
                maxleaf = std::max(maxleaf, cnt);
                maxd = std::max(maxd, it.depth + 1);
                sah += SAH_C_TRI * cnt * areas[c];
            }
            else
            {
                sah += SAH_C_NODE * areas[c];
                stack.push_back({refs[c], it.depth + 1});
            }
        }
    }
    for (int32_t q = 0; q < n; q++)
        if (!tri_seen[q])
            fail(MIPT_ERR_HIP, "BVH validation failed: triangle in no leaf");
    // Depth-first layout of the reachable interior nodes: a node's first
    // child follows it, so a traversal step into it often reads the 128-byte
    // line its parent was in. Hits do not depend on the node order.
    {
        std::vector<int32_t> remap(ninner, -1);
        std::vector<int32_t> order;
        order.reserve(inner);
        std::vector<int32_t> dfs{0};
        while (!dfs.empty())
        {
            const int32_t nd = dfs.back();
            dfs.pop_back();
            remap[nd] = (int32_t)order.size();
            order.push_back(nd);
            const int32_t c1 = hn[nd].d.y, c0 = hn[nd].d.x;
            if (n > 1 && c1 >= 0)
                dfs.push_back(c1);
            if (c0 >= 0)
                dfs.push_back(c0);
        }
        std::vector<BvhNode> laid(order.size());
        for (size_t i = 0; i < order.size(); i++)
        {
            BvhNode nd = hn[order[i]];
            if (nd.d.x >= 0)
                nd.d.x = remap[nd.d.x];
            if (n > 1 && nd.d.y >= 0)
                nd.d.y = remap[nd.d.y];
            laid[i] = nd;
        }
        HIP_CHECK(hipMemcpy(bvh.nodes.ptr, laid.data(), sizeof(BvhNode) * laid.size(), hipMemcpyHostToDevice));
    }
    bvh.node_count = inner;
    bvh.leaf_count = leaves;
    bvh.max_depth = maxd;
    bvh.max_leaf_size = maxleaf;
    bvh.sah_cost = root_area > 0 ? sah / root_area : 0.0;
    bvh.valid = true;
}

void build_lbvh(Ctx& ctx)
{
    const int32_t n = ctx.triangle_count;
    Bvh& bvh = ctx.bvh;
    bvh.valid = false;
    bvh.builder = MIPT_BVH_LBVH;
    bvh.triangle_count = n;
    if (n <= 0)
    {
        bvh.node_count = bvh.leaf_count = bvh.max_depth = bvh.max_leaf_size = 0;
        bvh.valid = true;
        return;
    }
    const float* verts = ctx.vertices.as<float>();
    const int32_t* idx = ctx.indices.as<int32_t>();

    hipEvent_t ev0, ev1;
    HIP_CHECK(hipEventCreate(&ev0));
    HIP_CHECK(hipEventCreate(&ev1));

    DeviceBuffer plo, phi, bounds, pad, partial, keys, keys_sorted, vals, vals_sorted, sort_tmp;
    const unsigned bgrid = bvhc::bounds_grid(n);
    plo.reserve(sizeof(float4) * n);
    phi.reserve(sizeof(float4) * n);
    bounds.reserve(sizeof(uint32_t) * 12);
    pad.reserve(sizeof(float));
    partial.reserve(sizeof(float) * 12 * bgrid);
    keys.reserve(sizeof(uint64_t) * n);
    keys_sorted.reserve(sizeof(uint64_t) * n);
    vals.reserve(sizeof(uint32_t) * n);
    vals_sorted.reserve(sizeof(uint32_t) * n);

    const hipStream_t s = nullptr;
    HIP_CHECK(hipEventRecord(ev0, s));
    const int B = 256;
    hipLaunchKernelGGL(bvhc::k_tri_boxes, dim3(bgrid), dim3(bvhc::BOUNDS_BLOCK), 0, s, verts, idx, n,
                       plo.as<float4>(), phi.as<float4>(), partial.as<float>());
    hipLaunchKernelGGL(bvhc::k_bounds_final, dim3(1), dim3(64), 0, s, partial.as<float>(), (int32_t)bgrid,
                       bounds.as<uint32_t>(), pad.as<float>());
    hipLaunchKernelGGL(k_morton, dim3(blocks(n, B)), dim3(B), 0, s, plo.as<float4>(), phi.as<float4>(), n,
                       bounds.as<uint32_t>(), keys.as<uint64_t>(), vals.as<uint32_t>());
    HIP_CHECK(hipGetLastError());

    size_t tmp_bytes = 0;
    HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys.as<uint64_t>(), keys_sorted.as<uint64_t>(),
                                        vals.as<uint32_t>(), vals_sorted.as<uint32_t>(), (size_t)n, 0u, 63u, s));
    sort_tmp.reserve(tmp_bytes > 0 ? tmp_bytes : 1);
    HIP_CHECK(rocprim::radix_sort_pairs(sort_tmp.ptr, tmp_bytes, keys.as<uint64_t>(), keys_sorted.as<uint64_t>(),
                                        vals.as<uint32_t>(), vals_sorted.as<uint32_t>(), (size_t)n, 0u, 63u, s));

    const int32_t ninner = n > 1 ? n - 1 : 1;
    DeviceBuffer child, node_parent, leaf_parent, range_lo, range_hi, visits, nlo, nhi, cost, collapse;
    child.reserve(sizeof(int32_t) * 2 * ninner);
    node_parent.reserve(sizeof(int32_t) * ninner);
    leaf_parent.reserve(sizeof(int32_t) * n);
    range_lo.reserve(sizeof(int32_t) * ninner);
    range_hi.reserve(sizeof(int32_t) * ninner);
    visits.reserve(sizeof(uint32_t) * ninner);
    nlo.reserve(sizeof(float4) * ninner);
    nhi.reserve(sizeof(float4) * ninner);
    cost.reserve(sizeof(float) * ninner);
    collapse.reserve(sizeof(uint8_t) * ninner);
    HIP_CHECK(hipMemsetAsync(visits.ptr, 0, sizeof(uint32_t) * ninner, s));
    HIP_CHECK(hipMemsetAsync(node_parent.ptr, 0xff, sizeof(int32_t) * ninner, s));
    HIP_CHECK(hipMemsetAsync(leaf_parent.ptr, 0xff, sizeof(int32_t) * n, s));
    HIP_CHECK(hipMemsetAsync(collapse.ptr, 0, sizeof(uint8_t) * ninner, s));

    bvh.nodes.reserve(sizeof(BvhNode) * ninner);
    bvh.tris.reserve(sizeof(TriRecord) * n);

    if (n == 1)
        write_single_triangle_root(ctx, plo.as<float4>(), phi.as<float4>());
    else
    {
        hipLaunchKernelGGL(k_karras, dim3(blocks(n - 1, B)), dim3(B), 0, s, keys_sorted.as<uint64_t>(), n,
                           child.as<int32_t>(), node_parent.as<int32_t>(), leaf_parent.as<int32_t>(),
                           range_lo.as<int32_t>(), range_hi.as<int32_t>());
        hipLaunchKernelGGL(k_refit, dim3(blocks(n, B)), dim3(B), 0, s, n, plo.as<float4>(), phi.as<float4>(),
                           vals_sorted.as<uint32_t>(), child.as<int32_t>(), node_parent.as<int32_t>(),
                           leaf_parent.as<int32_t>(), range_lo.as<int32_t>(), range_hi.as<int32_t>(),
                           visits.as<uint32_t>(), nlo.as<float4>(), nhi.as<float4>(), cost.as<float>(),
                           collapse.as<uint8_t>(), pad.as<float>());
        hipLaunchKernelGGL(k_emit_nodes, dim3(blocks(n - 1, B)), dim3(B), 0, s, n, plo.as<float4>(),
                           phi.as<float4>(), vals_sorted.as<uint32_t>(), child.as<int32_t>(), range_lo.as<int32_t>(),
                           range_hi.as<int32_t>(), nlo.as<float4>(), nhi.as<float4>(), collapse.as<uint8_t>(),
                           pad.as<float>(), bvh.nodes.as<BvhNode>());
    }
    hipLaunchKernelGGL(bvhc::k_emit_tris, dim3(blocks(n, B)), dim3(B), 0, s, verts, idx, n,
                       vals_sorted.as<uint32_t>(), bvh.tris.as<TriRecord>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(ev1, s));
    HIP_CHECK(hipEventSynchronize(ev1));
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
    bvh.build_ms = ms;
    HIP_CHECK(hipEventDestroy(ev0));
    HIP_CHECK(hipEventDestroy(ev1));
    // Statistics, validation and the depth-first layout: not part of the
    // build time.
    finalize_bvh(ctx, ninner);
}

} // namespace mipt
