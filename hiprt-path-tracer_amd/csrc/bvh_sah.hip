// bvh_sah.hip -- GPU top-down binned-SAH BVH builder (MIPT_BVH_SAH_BINNED).
//
// Second builder behind mipt_build_bvh, beside the LBVH: it replaces the
// high-quality mode of hiprtBuildGeometry (hiprtBuildFlagBitPreferHighQualityBuild,
// src/HIPRT-Orochi/HIPRTScene.h:87-150), whose implementation lives in the
// absent HIPRT library. The tree is built breadth first, one level of tasks
// per round; a task is a contiguous range of the primitive-index array `ref`:
//
//   * small tasks (<= SMALL_MAX primitives): one 256-thread block per task
//     loads the range into registers, reduces its box and centroid bounds,
//     bins the centroids into NB bins per axis in LDS, sweeps the 3 x (NB-1)
//     split planes (one wave per axis, scans by lane shuffles), decides leaf
//     or split by the SAH, partitions the range in place and appends the
//     two children to the next round's list;
//   * large tasks: the same steps spread over chunks of SMALL_MAX primitives
//     (one block per chunk; bins merged with global atomics, the partition
//     through per-chunk offsets computed on the host, so the result does
//     not depend on block timing).
//
// SAH (C_node = C_tri = 1): a task becomes a leaf when it has at most
// MAX_LEAF_SIZE triangles and count <= 1 + (A_L N_L + A_R N_R) / A. A task
// whose centroids all fall into one bin on every axis, or that lies deeper
// than DEPTH_LIMIT, is split at the median of its range (bounded depth,
// every side non-empty). The node boxes are unions of the padded triangle
// boxes (bvh_common.h), so they are conservative for the exact triangle
// test. Node ids come from an atomic counter; finalize_bvh then lays the
// reachable tree out depth first, so the emitted node array and the leaf
// order of the triangles are the same on every build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "bvh_common.h"

namespace mipt {
namespace {

using bvhc::f2o;
using bvhc::half_area;
using bvhc::o2f;

constexpr int NB = 32;           // bins per axis
constexpr int SB = 256;          // threads of a task block
constexpr int PER = 8;           // primitives per thread
constexpr int SMALL_MAX = SB * PER;
constexpr int DEPTH_LIMIT = 48;  // deeper tasks are split at the median
constexpr int32_t PENDING = 0x7fffffff;

struct SahTask {
    int32_t begin, count;
    int32_t slot;  // parent node * 2 + side; -1 for the root
    int32_t depth;
};

struct Chunk {
    int32_t task;   // index into this round's large-task list
    int32_t start;  // first primitive (absolute index into ref)
    int32_t end;
    int32_t left_off; // left primitives of the task's earlier chunks
};

struct LargeSplit {
    int32_t mode; // 1: bins, 2: median
    int32_t axis;
    int32_t k;    // last bin of the left side
    int32_t nl;   // primitives on the left
};

__device__ __forceinline__ float4 centroid(float4 lo, float4 hi)
{
    return make_float4(0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z), 0.0f);
}

__device__ __forceinline__ float comp(float4 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// Bin of centroid coordinate c on an axis with centroid minimum cmin and
// scale NB / extent. Every kernel bins with this one expression.
__device__ __forceinline__ int bin_of(float c, float cmin, float scale)
{
    const int b = (int)((c - cmin) * scale);
    return b < 0 ? 0 : (b > NB - 1 ? NB - 1 : b);
}

__device__ __forceinline__ float axis_scale(float cmin, float cmax)
{
    const float ext = cmax - cmin;
    return ext > 0.0f ? (float)NB / ext : 0.0f;
}

// 12 bounds (box lo3 hi3, centroid lo3 hi3) reduced over a block of SB
// threads; the result is in red_out[12] for every thread after the call.
__device__ __forceinline__ void block_bounds(float v[12], float (*red)[12], float* red_out)
{
    for (int k = 0; k < 12; k++)
        v[k] = ((k % 6) < 3) ? bvhc::wave_min(v[k]) : bvhc::wave_max(v[k]);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 12; k++) red[threadIdx.x / 64][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 12)
    {
        const int k = threadIdx.x;
        float r = red[0][k];
        for (int w = 1; w < SB / 64; w++) r = (k % 6) < 3 ? fminf(r, red[w][k]) : fmaxf(r, red[w][k]);
        red_out[k] = r;
    }
    __syncthreads();
}

__device__ __forceinline__ void bounds_init(float v[12])
{
    for (int k = 0; k < 12; k++) v[k] = (k % 6) < 3 ? INFINITY : -INFINITY;
}

__device__ __forceinline__ void bounds_add(float v[12], float4 lo, float4 hi, float4 c)
{
    v[0] = fminf(v[0], lo.x); v[1] = fminf(v[1], lo.y); v[2] = fminf(v[2], lo.z);
    v[3] = fmaxf(v[3], hi.x); v[4] = fmaxf(v[4], hi.y); v[5] = fmaxf(v[5], hi.z);
    v[6] = fminf(v[6], c.x); v[7] = fminf(v[7], c.y); v[8] = fminf(v[8], c.z);
    v[9] = fmaxf(v[9], c.x); v[10] = fmaxf(v[10], c.y); v[11] = fmaxf(v[11], c.z);
}

// Bin record: count, box lo (3, ordered uint), box hi (3), unused.
constexpr int BIN_WORDS = 8;

__device__ __forceinline__ void bins_clear(uint32_t* bins, int words)
{
    for (int i = threadIdx.x; i < words; i += blockDim.x)
    {
        const int f = i % BIN_WORDS;
        bins[i] = f == 0 ? 0u : (f <= 3 ? f2o(INFINITY) : (f <= 6 ? f2o(-INFINITY) : 0u));
    }
}

__device__ __forceinline__ void bin_add(uint32_t* bin, float4 lo, float4 hi)
{
    atomicAdd(&bin[0], 1u);
    atomicMin(&bin[1], f2o(lo.x));
    atomicMin(&bin[2], f2o(lo.y));
    atomicMin(&bin[3], f2o(lo.z));
    atomicMax(&bin[4], f2o(hi.x));
    atomicMax(&bin[5], f2o(hi.y));
    atomicMax(&bin[6], f2o(hi.z));
}

struct AxisBest {
    float cost; // A_L N_L + A_R N_R; +inf when no plane splits the task
    int32_t k, nl;
    float4 llo, lhi, rlo, rhi;
};

// Sweep of one axis by one wave: lane l < NB holds bin l of `bins`
// (NB x BIN_WORDS words). Returns the best plane (smallest cost, ties to the
// lowest plane) with both sides' boxes, the same in every lane.
__device__ __forceinline__ AxisBest sweep_axis(const uint32_t* bins, bool axis_ok)
{
    const int lane = threadIdx.x & 63;
    uint32_t cnt = 0;
    float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
    float4 hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
    if (lane < NB && axis_ok)
    {
        const uint32_t* b = bins + lane * BIN_WORDS;
        cnt = b[0];
        if (cnt)
        {
            lo = make_float4(o2f(b[1]), o2f(b[2]), o2f(b[3]), 0.0f);
            hi = make_float4(o2f(b[4]), o2f(b[5]), o2f(b[6]), 0.0f);
        }
    }
    // inclusive prefix (lanes <= l) and suffix (lanes >= l)
    uint32_t pc = cnt, sc = cnt;
    float4 plo = lo, phi = hi, slo = lo, shi = hi;
    for (int off = 1; off < 64; off <<= 1)
    {
        const uint32_t uc = __shfl_up(pc, off, 64);
        const float ulx = __shfl_up(plo.x, off, 64), uly = __shfl_up(plo.y, off, 64), ulz = __shfl_up(plo.z, off, 64);
        const float uhx = __shfl_up(phi.x, off, 64), uhy = __shfl_up(phi.y, off, 64), uhz = __shfl_up(phi.z, off, 64);
        const uint32_t dc = __shfl_down(sc, off, 64);
        const float dlx = __shfl_down(slo.x, off, 64), dly = __shfl_down(slo.y, off, 64),
                    dlz = __shfl_down(slo.z, off, 64);
        const float dhx = __shfl_down(shi.x, off, 64), dhy = __shfl_down(shi.y, off, 64),
                    dhz = __shfl_down(shi.z, off, 64);
        if (lane >= off)
        {
            pc += uc;
            plo = make_float4(fminf(plo.x, ulx), fminf(plo.y, uly), fminf(plo.z, ulz), 0.0f);
            phi = make_float4(fmaxf(phi.x, uhx), fmaxf(phi.y, uhy), fmaxf(phi.z, uhz), 0.0f);
        }
        if (lane + off < 64)
        {
            sc += dc;
            slo = make_float4(fminf(slo.x, dlx), fminf(slo.y, dly), fminf(slo.z, dlz), 0.0f);
            shi = make_float4(fmaxf(shi.x, dhx), fmaxf(shi.y, dhy), fmaxf(shi.z, dhz), 0.0f);
        }
    }
    // plane after bin `lane`: left = prefix[lane], right = suffix[lane + 1]
    const uint32_t nr = __shfl_down(sc, 1, 64);
    const float4 rlo = make_float4(__shfl_down(slo.x, 1, 64), __shfl_down(slo.y, 1, 64), __shfl_down(slo.z, 1, 64), 0.0f);
    const float4 rhi = make_float4(__shfl_down(shi.x, 1, 64), __shfl_down(shi.y, 1, 64), __shfl_down(shi.z, 1, 64), 0.0f);
    float cost = INFINITY;
    if (axis_ok && lane < NB - 1 && pc > 0 && nr > 0)
        cost = half_area(plo, phi) * (float)pc + half_area(rlo, rhi) * (float)nr;
    // argmin over the wave, ties to the lower plane
    float bc = cost;
    int bk = lane;
    for (int off = 32; off > 0; off >>= 1)
    {
        const float oc = __shfl_xor(bc, off, 64);
        const int ok = __shfl_xor(bk, off, 64);
        if (oc < bc || (oc == bc && ok < bk))
        {
            bc = oc;
            bk = ok;
        }
    }
    AxisBest r;
    r.cost = bc;
    r.k = bk;
    r.nl = (int32_t)__shfl(pc, bk, 64);
    r.llo = make_float4(__shfl(plo.x, bk, 64), __shfl(plo.y, bk, 64), __shfl(plo.z, bk, 64), 0.0f);
    r.lhi = make_float4(__shfl(phi.x, bk, 64), __shfl(phi.y, bk, 64), __shfl(phi.z, bk, 64), 0.0f);
    r.rlo = make_float4(__shfl(rlo.x, bk, 64), __shfl(rlo.y, bk, 64), __shfl(rlo.z, bk, 64), 0.0f);
    r.rhi = make_float4(__shfl(rhi.x, bk, 64), __shfl(rhi.y, bk, 64), __shfl(rhi.z, bk, 64), 0.0f);
    return r;
}

__device__ __forceinline__ void write_node_boxes(BvhNode* node, float4 llo, float4 lhi, float4 rlo, float4 rhi,
                                                 int32_t ref0, int32_t ref1)
{
    BvhNode nd;
    nd.a = make_float4(llo.x, llo.y, llo.z, lhi.x);
    nd.b = make_float4(lhi.y, lhi.z, rlo.x, rlo.y);
    nd.c = make_float4(rlo.z, rhi.x, rhi.y, rhi.z);
    nd.d = make_int4(ref0, ref1, 0, 0);
    *node = nd;
}

__device__ __forceinline__ void set_slot(BvhNode* nodes, int32_t slot, int32_t ref)
{
    if (slot < 0)
        return;
    int32_t* d = reinterpret_cast<int32_t*>(&nodes[slot >> 1].d);
    d[slot & 1] = ref;
}

// Appends child (begin, count) of `node` (side s): a single triangle is a
// leaf at once, other children go to the next round.
__device__ __forceinline__ void push_child(BvhNode* nodes, int32_t node, int s, int32_t begin, int32_t count,
                                           int32_t depth, uint32_t* counters, SahTask* next_small,
                                           SahTask* next_large)
{
    if (count == 1)
        return; // written with the node record
    SahTask t{begin, count, node * 2 + s, depth};
    if (count > SMALL_MAX)
        next_large[atomicAdd(&counters[2], 1u)] = t;
    else
        next_small[atomicAdd(&counters[1], 1u)] = t;
}

// Exclusive scan of one value per thread over a block of SB threads; `tot`
// receives the block total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum, uint32_t& tot)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x / 64;
    uint32_t inc = v;
    for (int off = 1; off < 64; off <<= 1)
    {
        const uint32_t u = __shfl_up(inc, off, 64);
        if (lane >= off)
            inc += u;
    }
    if (lane == 63)
        wsum[w] = inc;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (int i = 0; i < SB / 64; i++)
    {
        if (i < w)
            base += wsum[i];
        all += wsum[i];
    }
    tot = all;
    __syncthreads();
    return base + inc - v;
}

// One round of every small task: one block each.
__global__ __launch_bounds__(SB) void k_sah_small(const SahTask* __restrict__ tasks, uint32_t* __restrict__ ref,
                                                  const float4* __restrict__ blo, const float4* __restrict__ bhi,
                                                  BvhNode* __restrict__ nodes, uint32_t* __restrict__ counters,
                                                  SahTask* __restrict__ next_small, SahTask* __restrict__ next_large)
{
    __shared__ uint32_t bins[3 * NB * BIN_WORDS];
    __shared__ float red[SB / 64][12];
    __shared__ float tb[12];
    __shared__ AxisBest best[3];
    __shared__ uint32_t wsum[SB / 64];
    __shared__ float hb[12]; // median split: left box lo/hi, right box lo/hi
    const SahTask task = tasks[blockIdx.x];
    const int t = threadIdx.x;

    uint32_t r[PER];
    float4 c[PER];
    float v[12];
    bounds_init(v);
#pragma unroll
    for (int j = 0; j < PER; j++)
    {
        const int i = t * PER + j;
        r[j] = 0u;
        c[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (i < task.count)
        {
            r[j] = ref[task.begin + i];
            const float4 lo = blo[r[j]], hi = bhi[r[j]];
            c[j] = centroid(lo, hi);
            bounds_add(v, lo, hi, c[j]);
        }
    }
    block_bounds(v, red, tb);
    const float4 tlo = make_float4(tb[0], tb[1], tb[2], 0.0f), thi = make_float4(tb[3], tb[4], tb[5], 0.0f);
    float cmin[3], scale[3];
    for (int a = 0; a < 3; a++)
    {
        cmin[a] = tb[6 + a];
        scale[a] = axis_scale(tb[6 + a], tb[9 + a]);
    }

    // binning
    bins_clear(bins, 3 * NB * BIN_WORDS);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; j++)
    {
        if (t * PER + j < task.count)
        {
            const float4 lo = blo[r[j]], hi = bhi[r[j]];
            for (int a = 0; a < 3; a++)
                if (scale[a] > 0.0f)
                    bin_add(bins + (a * NB + bin_of(comp(c[j], a), cmin[a], scale[a])) * BIN_WORDS, lo, hi);
        }
    }
    __syncthreads();
    if (t < 3 * 64)
    {
        const int a = t / 64;
        const AxisBest b = sweep_axis(bins + a * NB * BIN_WORDS, scale[a] > 0.0f);
        if ((t & 63) == 0)
            best[a] = b;
    }
    __syncthreads();
    int axis = 0;
    for (int a = 1; a < 3; a++)
        if (best[a].cost < best[axis].cost)
            axis = a;
    const AxisBest b = best[axis];
    const float area = half_area(tlo, thi);
    const float split_cost = bvhc::SAH_C_NODE + bvhc::SAH_C_TRI * b.cost / area;
    const float leaf_cost = bvhc::SAH_C_TRI * (float)task.count;
    const bool can_leaf = task.count <= MAX_LEAF_SIZE && task.slot >= 0;
    if (can_leaf && (leaf_cost <= split_cost || task.depth >= DEPTH_LIMIT || !(b.cost < INFINITY)))
    {
        if (t == 0)
            set_slot(nodes, task.slot, make_leaf(task.begin, task.count));
        return;
    }
    const bool median = !(b.cost < INFINITY) || task.depth >= DEPTH_LIMIT;
    int32_t nl;
    float4 llo, lhi, rlo, rhi;
    if (median)
    {
        // the range is split in place at its middle: no primitive moves
        nl = task.count / 2;
        float w[12];
        bounds_init(w); // lo3 hi3 of the left half, lo3 hi3 of the right half (in the centroid slots' order)
#pragma unroll
        for (int j = 0; j < PER; j++)
        {
            const int i = t * PER + j;
            if (i >= task.count)
                continue;
            const float4 lo = blo[r[j]], hi = bhi[r[j]];
            const int o = i < nl ? 0 : 6;
            w[o + 0] = fminf(w[o + 0], lo.x); w[o + 1] = fminf(w[o + 1], lo.y); w[o + 2] = fminf(w[o + 2], lo.z);
            w[o + 3] = fmaxf(w[o + 3], hi.x); w[o + 4] = fmaxf(w[o + 4], hi.y); w[o + 5] = fmaxf(w[o + 5], hi.z);
        }
        block_bounds(w, red, hb);
        llo = make_float4(hb[0], hb[1], hb[2], 0.0f);
        lhi = make_float4(hb[3], hb[4], hb[5], 0.0f);
        rlo = make_float4(hb[6], hb[7], hb[8], 0.0f);
        rhi = make_float4(hb[9], hb[10], hb[11], 0.0f);
    }
    else
    {
        nl = b.nl;
        llo = b.llo;
        lhi = b.lhi;
        rlo = b.rlo;
        rhi = b.rhi;
        // partition in place: every read of the range is in registers
        uint32_t flags = 0u, lc = 0u;
#pragma unroll
        for (int j = 0; j < PER; j++)
        {
            if (t * PER + j < task.count && bin_of(comp(c[j], axis), cmin[axis], scale[axis]) <= b.k)
            {
                flags |= 1u << j;
                lc++;
            }
        }
        uint32_t tot;
        const uint32_t lbase = block_excl_scan(lc, wsum, tot);
        const int before = min(t * PER, task.count);
        uint32_t lp = lbase, rp = (uint32_t)nl + (uint32_t)before - lbase;
#pragma unroll
        for (int j = 0; j < PER; j++)
        {
            if (t * PER + j >= task.count)
                continue;
            if (flags & (1u << j))
                ref[task.begin + lp++] = r[j];
            else
                ref[task.begin + rp++] = r[j];
        }
    }
    if (t == 0)
    {
        const int32_t node = (int32_t)atomicAdd(&counters[0], 1u);
        const int32_t nr = task.count - nl;
        write_node_boxes(nodes + node, llo, lhi, rlo, rhi, nl == 1 ? make_leaf(task.begin, 1) : PENDING,
                         nr == 1 ? make_leaf(task.begin + nl, 1) : PENDING);
        set_slot(nodes, task.slot, node);
        push_child(nodes, node, 0, task.begin, nl, task.depth + 1, counters, next_small, next_large);
        push_child(nodes, node, 1, task.begin + nl, nr, task.depth + 1, counters, next_small, next_large);
    }
}

// ---- large tasks ------------------------------------------------------
// Per large task: acc[12] (ordered-uint box and centroid bounds) and
// bins[3 * NB * BIN_WORDS].
constexpr int LARGE_WORDS = 12 + 3 * NB * BIN_WORDS;

__global__ void k_large_init(uint32_t* __restrict__ acc, int32_t ntasks)
{
    uint32_t* a = acc + (size_t)blockIdx.x * LARGE_WORDS;
    if (blockIdx.x >= (unsigned)ntasks)
        return;
    for (int i = threadIdx.x; i < 12; i += blockDim.x)
        a[i] = (i % 6) < 3 ? f2o(INFINITY) : f2o(-INFINITY);
    uint32_t* b = a + 12;
    for (int i = threadIdx.x; i < 3 * NB * BIN_WORDS; i += blockDim.x)
    {
        const int f = i % BIN_WORDS;
        b[i] = f == 0 ? 0u : (f <= 3 ? f2o(INFINITY) : (f <= 6 ? f2o(-INFINITY) : 0u));
    }
}

__global__ __launch_bounds__(SB) void k_large_bounds(const Chunk* __restrict__ chunks, const uint32_t* __restrict__ ref,
                                                     const float4* __restrict__ blo, const float4* __restrict__ bhi,
                                                     uint32_t* __restrict__ acc)
{
    __shared__ float red[SB / 64][12];
    __shared__ float out[12];
    const Chunk ch = chunks[blockIdx.x];
    float v[12];
    bounds_init(v);
    for (int j = 0; j < PER; j++)
    {
        const int i = ch.start + threadIdx.x * PER + j;
        if (i < ch.end)
        {
            const uint32_t p = ref[i];
            const float4 lo = blo[p], hi = bhi[p];
            bounds_add(v, lo, hi, centroid(lo, hi));
        }
    }
    block_bounds(v, red, out);
    if (threadIdx.x < 12)
    {
        uint32_t* a = acc + (size_t)ch.task * LARGE_WORDS;
        const int k = threadIdx.x;
        if ((k % 6) < 3)
            atomicMin(&a[k], f2o(out[k]));
        else
            atomicMax(&a[k], f2o(out[k]));
    }
}

__device__ __forceinline__ void large_axes(const uint32_t* a, float cmin[3], float scale[3])
{
    for (int k = 0; k < 3; k++)
    {
        cmin[k] = o2f(a[6 + k]);
        scale[k] = axis_scale(cmin[k], o2f(a[9 + k]));
    }
}

__global__ __launch_bounds__(SB) void k_large_bin(const Chunk* __restrict__ chunks, const uint32_t* __restrict__ ref,
                                                  const float4* __restrict__ blo, const float4* __restrict__ bhi,
                                                  uint32_t* __restrict__ acc)
{
    __shared__ uint32_t bins[3 * NB * BIN_WORDS];
    const Chunk ch = chunks[blockIdx.x];
    uint32_t* a = acc + (size_t)ch.task * LARGE_WORDS;
    float cmin[3], scale[3];
    large_axes(a, cmin, scale);
    bins_clear(bins, 3 * NB * BIN_WORDS);
    __syncthreads();
    for (int j = 0; j < PER; j++)
    {
        const int i = ch.start + threadIdx.x * PER + j;
        if (i >= ch.end)
            break;
        const uint32_t p = ref[i];
        const float4 lo = blo[p], hi = bhi[p];
        const float4 c = centroid(lo, hi);
        for (int ax = 0; ax < 3; ax++)
            if (scale[ax] > 0.0f)
                bin_add(bins + (ax * NB + bin_of(comp(c, ax), cmin[ax], scale[ax])) * BIN_WORDS, lo, hi);
    }
    __syncthreads();
    uint32_t* g = a + 12;
    for (int i = threadIdx.x; i < 3 * NB; i += blockDim.x)
    {
        const uint32_t* b = bins + i * BIN_WORDS;
        if (b[0] == 0u)
            continue;
        uint32_t* d = g + i * BIN_WORDS;
        atomicAdd(&d[0], b[0]);
        for (int f = 1; f <= 3; f++) atomicMin(&d[f], b[f]);
        for (int f = 4; f <= 6; f++) atomicMax(&d[f], b[f]);
    }
}

// One block of 3 waves per large task: the split, the node and its boxes.
__global__ __launch_bounds__(192) void k_large_split(const SahTask* __restrict__ tasks,
                                                     const uint32_t* __restrict__ acc, BvhNode* __restrict__ nodes,
                                                     uint32_t* __restrict__ counters, LargeSplit* __restrict__ split)
{
    __shared__ AxisBest best[3];
    const SahTask task = tasks[blockIdx.x];
    const uint32_t* a = acc + (size_t)blockIdx.x * LARGE_WORDS;
    float cmin[3], scale[3];
    large_axes(a, cmin, scale);
    const int ax = threadIdx.x / 64;
    const AxisBest b = sweep_axis(a + 12 + ax * NB * BIN_WORDS, scale[ax] > 0.0f);
    if ((threadIdx.x & 63) == 0)
        best[ax] = b;
    __syncthreads();
    if (threadIdx.x != 0)
        return;
    int axis = 0;
    for (int k = 1; k < 3; k++)
        if (best[k].cost < best[axis].cost)
            axis = k;
    const AxisBest& bb = best[axis];
    LargeSplit s;
    float4 llo, lhi, rlo, rhi;
    if (!(bb.cost < INFINITY) || task.depth >= DEPTH_LIMIT)
    {
        // median of the range; both children keep the task's box
        s.mode = 2;
        s.axis = 0;
        s.k = 0;
        s.nl = task.count / 2;
        llo = rlo = make_float4(o2f(a[0]), o2f(a[1]), o2f(a[2]), 0.0f);
        lhi = rhi = make_float4(o2f(a[3]), o2f(a[4]), o2f(a[5]), 0.0f);
    }
    else
    {
        s.mode = 1;
        s.axis = axis;
        s.k = bb.k;
        s.nl = bb.nl;
        llo = bb.llo;
        lhi = bb.lhi;
        rlo = bb.rlo;
        rhi = bb.rhi;
    }
    const int32_t node = (int32_t)atomicAdd(&counters[0], 1u);
    const int32_t nr = task.count - s.nl;
    write_node_boxes(nodes + node, llo, lhi, rlo, rhi, s.nl == 1 ? make_leaf(task.begin, 1) : PENDING,
                     nr == 1 ? make_leaf(task.begin + s.nl, 1) : PENDING);
    set_slot(nodes, task.slot, node);
    split[blockIdx.x] = s;
    // the node id travels to k_large_children through the split record's slot
    counters[8 + blockIdx.x] = (uint32_t)node;
}

__device__ __forceinline__ bool goes_left(const LargeSplit& s, const SahTask& task, int32_t i, float4 c,
                                          const float cmin[3], const float scale[3])
{
    if (s.mode == 2)
        return i - task.begin < s.nl;
    return bin_of(comp(c, s.axis), cmin[s.axis], scale[s.axis]) <= s.k;
}

__global__ __launch_bounds__(SB) void k_large_count(const Chunk* __restrict__ chunks, const SahTask* __restrict__ tasks,
                                                    const uint32_t* __restrict__ ref, const float4* __restrict__ blo,
                                                    const float4* __restrict__ bhi, const uint32_t* __restrict__ acc,
                                                    const LargeSplit* __restrict__ split,
                                                    uint32_t* __restrict__ chunk_left)
{
    __shared__ uint32_t wsum[SB / 64];
    const Chunk ch = chunks[blockIdx.x];
    const SahTask task = tasks[ch.task];
    const LargeSplit s = split[ch.task];
    float cmin[3], scale[3];
    large_axes(acc + (size_t)ch.task * LARGE_WORDS, cmin, scale);
    uint32_t lc = 0;
    for (int j = 0; j < PER; j++)
    {
        const int i = ch.start + threadIdx.x * PER + j;
        if (i >= ch.end)
            break;
        const uint32_t p = ref[i];
        if (goes_left(s, task, i, centroid(blo[p], bhi[p]), cmin, scale))
            lc++;
    }
    uint32_t tot;
    block_excl_scan(lc, wsum, tot);
    if (threadIdx.x == 0)
        chunk_left[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SB) void k_large_scatter(const Chunk* __restrict__ chunks,
                                                      const SahTask* __restrict__ tasks,
                                                      const uint32_t* __restrict__ ref, const float4* __restrict__ blo,
                                                      const float4* __restrict__ bhi, const uint32_t* __restrict__ acc,
                                                      const LargeSplit* __restrict__ split,
                                                      uint32_t* __restrict__ ref_out)
{
    __shared__ uint32_t wsum[SB / 64];
    const Chunk ch = chunks[blockIdx.x];
    const SahTask task = tasks[ch.task];
    const LargeSplit s = split[ch.task];
    float cmin[3], scale[3];
    large_axes(acc + (size_t)ch.task * LARGE_WORDS, cmin, scale);
    uint32_t r[PER];
    uint32_t flags = 0u, lc = 0u;
#pragma unroll
    for (int j = 0; j < PER; j++)
    {
        const int i = ch.start + (int)threadIdx.x * PER + j;
        r[j] = 0u;
        if (i < ch.end)
        {
            r[j] = ref[i];
            if (goes_left(s, task, i, centroid(blo[r[j]], bhi[r[j]]), cmin, scale))
            {
                flags |= 1u << j;
                lc++;
            }
        }
    }
    uint32_t tot;
    const uint32_t lbase = block_excl_scan(lc, wsum, tot);
    const int32_t in_task = ch.start - task.begin;            // primitives of earlier chunks
    const int32_t before = min((int)threadIdx.x * PER, ch.end - ch.start); // of earlier threads
    uint32_t lp = (uint32_t)task.begin + (uint32_t)ch.left_off + lbase;
    uint32_t rp = (uint32_t)task.begin + (uint32_t)s.nl + (uint32_t)(in_task - ch.left_off) + (uint32_t)before - lbase;
#pragma unroll
    for (int j = 0; j < PER; j++)
    {
        if (ch.start + (int)threadIdx.x * PER + j >= ch.end)
            continue;
        if (flags & (1u << j))
            ref_out[lp++] = r[j];
        else
            ref_out[rp++] = r[j];
    }
}

__global__ void k_large_copy(const Chunk* __restrict__ chunks, const uint32_t* __restrict__ src,
                             uint32_t* __restrict__ dst)
{
    const Chunk ch = chunks[blockIdx.x];
    for (int i = ch.start + threadIdx.x; i < ch.end; i += blockDim.x)
        dst[i] = src[i];
}

__global__ void k_large_children(const SahTask* __restrict__ tasks, int32_t ntasks, const LargeSplit* __restrict__ split,
                                 BvhNode* __restrict__ nodes, uint32_t* __restrict__ counters,
                                 SahTask* __restrict__ next_small, SahTask* __restrict__ next_large)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntasks)
        return;
    const SahTask task = tasks[i];
    const LargeSplit s = split[i];
    const int32_t node = (int32_t)counters[8 + i];
    push_child(nodes, node, 0, task.begin, s.nl, task.depth + 1, counters, next_small, next_large);
    push_child(nodes, node, 1, task.begin + s.nl, task.count - s.nl, task.depth + 1, counters, next_small,
               next_large);
}

__global__ void k_pad_boxes(const float4* __restrict__ lo, const float4* __restrict__ hi, int32_t n,
                            const float* __restrict__ pad, float4* __restrict__ plo, float4* __restrict__ phi,
                            uint32_t* __restrict__ ref)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float4 a, b;
    bvhc::pad_box(lo[i], hi[i], pad[0], a, b);
    plo[i] = a;
    phi[i] = b;
    ref[i] = (uint32_t)i;
}

} // namespace

void build_sah_binned(Ctx& ctx)
{
    const int32_t n = ctx.triangle_count;
    Bvh& bvh = ctx.bvh;
    bvh.valid = false;
    bvh.builder = MIPT_BVH_SAH_BINNED;
    bvh.triangle_count = n;
    if (n <= 0)
    {
        bvh.node_count = bvh.leaf_count = bvh.max_depth = bvh.max_leaf_size = 0;
        bvh.valid = true;
        return;
    }
    const float* verts = ctx.vertices.as<float>();
    const int32_t* idx = ctx.indices.as<int32_t>();
    const hipStream_t s = nullptr;

    hipEvent_t ev0, ev1;
    HIP_CHECK(hipEventCreate(&ev0));
    HIP_CHECK(hipEventCreate(&ev1));

    const unsigned bgrid = bvhc::bounds_grid(n);
    DeviceBuffer lo, hi, plo, phi, partial, bounds, pad, ref, ref_tmp, counters;
    lo.reserve(sizeof(float4) * n);
    hi.reserve(sizeof(float4) * n);
    plo.reserve(sizeof(float4) * n);
    phi.reserve(sizeof(float4) * n);
    partial.reserve(sizeof(float) * 12 * bgrid);
    bounds.reserve(sizeof(uint32_t) * 12);
    pad.reserve(sizeof(float));
    ref.reserve(sizeof(uint32_t) * n);
    ref_tmp.reserve(sizeof(uint32_t) * n);
    // Task lists: a round holds disjoint ranges of >= 2 triangles.
    const size_t max_tasks = (size_t)n / 2 + 1;
    const size_t max_large = (size_t)n / (SMALL_MAX + 1) + 1;
    DeviceBuffer small[2], large[2], acc, lsplit, chunk_buf, chunk_left;
    small[0].reserve(sizeof(SahTask) * max_tasks);
    small[1].reserve(sizeof(SahTask) * max_tasks);
    large[0].reserve(sizeof(SahTask) * max_large);
    large[1].reserve(sizeof(SahTask) * max_large);
    acc.reserve(sizeof(uint32_t) * LARGE_WORDS * max_large);
    lsplit.reserve(sizeof(LargeSplit) * max_large);
    // counters: [0] nodes, [1] next small tasks, [2] next large tasks,
    // [8 + i] node of large task i
    counters.reserve(sizeof(uint32_t) * (8 + max_large));
    const size_t max_chunks = (size_t)n / SMALL_MAX + max_large + 1;
    chunk_buf.reserve(sizeof(Chunk) * max_chunks);
    chunk_left.reserve(sizeof(uint32_t) * max_chunks);

    const int32_t ninner = n > 1 ? n - 1 : 1;
    bvh.nodes.reserve(sizeof(BvhNode) * ninner);
    bvh.tris.reserve(sizeof(TriRecord) * n);

    HIP_CHECK(hipEventRecord(ev0, s));
    hipLaunchKernelGGL(bvhc::k_tri_boxes, dim3(bgrid), dim3(bvhc::BOUNDS_BLOCK), 0, s, verts, idx, n,
                       lo.as<float4>(), hi.as<float4>(), partial.as<float>());
    hipLaunchKernelGGL(bvhc::k_bounds_final, dim3(1), dim3(64), 0, s, partial.as<float>(), (int32_t)bgrid,
                       bounds.as<uint32_t>(), pad.as<float>());
    hipLaunchKernelGGL(k_pad_boxes, dim3(bvhc::blocks(n, 256)), dim3(256), 0, s, lo.as<float4>(), hi.as<float4>(), n,
                       pad.as<float>(), plo.as<float4>(), phi.as<float4>(), ref.as<uint32_t>());
    HIP_CHECK(hipGetLastError());

    if (n == 1)
        write_single_triangle_root(ctx, lo.as<float4>(), hi.as<float4>());
    else
    {
        HIP_CHECK(hipMemsetAsync(counters.ptr, 0, sizeof(uint32_t) * 8, s));
        const SahTask root{0, n, -1, 0};
        int cur = 0;
        uint32_t nsmall = 0, nlarge = 0;
        if (n > SMALL_MAX)
        {
            HIP_CHECK(hipMemcpyAsync(large[0].ptr, &root, sizeof(root), hipMemcpyHostToDevice, s));
            nlarge = 1;
        }
        else
        {
            HIP_CHECK(hipMemcpyAsync(small[0].ptr, &root, sizeof(root), hipMemcpyHostToDevice, s));
            nsmall = 1;
        }
        std::vector<SahTask> ht;
        std::vector<Chunk> hc;
        std::vector<uint32_t> hleft;
        std::vector<LargeSplit> hs;
        for (int round = 0; nsmall + nlarge > 0; round++)
        {
            if (round > 4 * 64 + 64)
                fail(MIPT_ERR_HIP, "SAH build: too many rounds");
            const int nxt = cur ^ 1;
            HIP_CHECK(hipMemsetAsync(counters.as<uint32_t>() + 1, 0, sizeof(uint32_t) * 2, s));
            if (nlarge > 0)
            {
                ht.resize(nlarge);
                HIP_CHECK(hipMemcpyAsync(ht.data(), large[cur].ptr, sizeof(SahTask) * nlarge, hipMemcpyDeviceToHost, s));
                HIP_CHECK(hipStreamSynchronize(s));
                hc.clear();
                for (uint32_t t = 0; t < nlarge; t++)
                    for (int32_t a = ht[t].begin; a < ht[t].begin + ht[t].count; a += SMALL_MAX)
                        hc.push_back(Chunk{(int32_t)t, a, std::min(a + SMALL_MAX, ht[t].begin + ht[t].count), 0});
                const uint32_t nch = (uint32_t)hc.size();
                if (nch > max_chunks)
                    fail(MIPT_ERR_HIP, "SAH build: chunk list overflow");
                HIP_CHECK(hipMemcpyAsync(chunk_buf.ptr, hc.data(), sizeof(Chunk) * nch, hipMemcpyHostToDevice, s));
                const SahTask* lt = large[cur].as<SahTask>();
                const Chunk* chk = chunk_buf.as<Chunk>();
                hipLaunchKernelGGL(k_large_init, dim3(nlarge), dim3(256), 0, s, acc.as<uint32_t>(), (int32_t)nlarge);
                hipLaunchKernelGGL(k_large_bounds, dim3(nch), dim3(SB), 0, s, chk, ref.as<uint32_t>(), plo.as<float4>(),
                                   phi.as<float4>(), acc.as<uint32_t>());
                hipLaunchKernelGGL(k_large_bin, dim3(nch), dim3(SB), 0, s, chk, ref.as<uint32_t>(), plo.as<float4>(),
                                   phi.as<float4>(), acc.as<uint32_t>());
                hipLaunchKernelGGL(k_large_split, dim3(nlarge), dim3(192), 0, s, lt, acc.as<uint32_t>(),
                                   bvh.nodes.as<BvhNode>(), counters.as<uint32_t>(), lsplit.as<LargeSplit>());
                hipLaunchKernelGGL(k_large_count, dim3(nch), dim3(SB), 0, s, chk, lt, ref.as<uint32_t>(),
                                   plo.as<float4>(), phi.as<float4>(), acc.as<uint32_t>(), lsplit.as<LargeSplit>(),
                                   chunk_left.as<uint32_t>());
                HIP_CHECK(hipGetLastError());
                hleft.resize(nch);
                hs.resize(nlarge);
                HIP_CHECK(hipMemcpyAsync(hleft.data(), chunk_left.ptr, sizeof(uint32_t) * nch, hipMemcpyDeviceToHost, s));
                HIP_CHECK(hipMemcpyAsync(hs.data(), lsplit.ptr, sizeof(LargeSplit) * nlarge, hipMemcpyDeviceToHost, s));
                HIP_CHECK(hipStreamSynchronize(s));
                // left offsets of each chunk within its task; the chunks'
                // left counts must add up to the split's count
                std::vector<int64_t> left_sum(nlarge, 0);
                for (uint32_t c = 0; c < nch; c++)
                {
                    hc[c].left_off = (int32_t)left_sum[hc[c].task];
                    left_sum[hc[c].task] += hleft[c];
                }
                for (uint32_t t = 0; t < nlarge; t++)
                    if (left_sum[t] != hs[t].nl || hs[t].nl < 1 || hs[t].nl >= ht[t].count)
                        fail(MIPT_ERR_HIP, "SAH build: partition counts disagree with the split");
                HIP_CHECK(hipMemcpyAsync(chunk_buf.ptr, hc.data(), sizeof(Chunk) * nch, hipMemcpyHostToDevice, s));
                hipLaunchKernelGGL(k_large_scatter, dim3(nch), dim3(SB), 0, s, chk, lt, ref.as<uint32_t>(),
                                   plo.as<float4>(), phi.as<float4>(), acc.as<uint32_t>(), lsplit.as<LargeSplit>(),
                                   ref_tmp.as<uint32_t>());
                hipLaunchKernelGGL(k_large_copy, dim3(nch), dim3(256), 0, s, chk, ref_tmp.as<uint32_t>(),
                                   ref.as<uint32_t>());
                hipLaunchKernelGGL(k_large_children, dim3(bvhc::blocks(nlarge, 64)), dim3(64), 0, s, lt,
                                   (int32_t)nlarge, lsplit.as<LargeSplit>(), bvh.nodes.as<BvhNode>(),
                                   counters.as<uint32_t>(), small[nxt].as<SahTask>(), large[nxt].as<SahTask>());
                HIP_CHECK(hipGetLastError());
            }
            if (nsmall > 0)
            {
                hipLaunchKernelGGL(k_sah_small, dim3(nsmall), dim3(SB), 0, s, small[cur].as<SahTask>(),
                                   ref.as<uint32_t>(), plo.as<float4>(), phi.as<float4>(), bvh.nodes.as<BvhNode>(),
                                   counters.as<uint32_t>(), small[nxt].as<SahTask>(), large[nxt].as<SahTask>());
                HIP_CHECK(hipGetLastError());
            }
            uint32_t hcnt[3];
            HIP_CHECK(hipMemcpyAsync(hcnt, counters.ptr, sizeof(hcnt), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            if (hcnt[0] > (uint32_t)ninner || hcnt[1] > max_tasks || hcnt[2] > max_large)
                fail(MIPT_ERR_HIP, "SAH build: node or task capacity exceeded");
            nsmall = hcnt[1];
            nlarge = hcnt[2];
            cur = nxt;
        }
    }
    hipLaunchKernelGGL(bvhc::k_emit_tris, dim3(bvhc::blocks(n, 256)), dim3(256), 0, s, verts, idx, n,
                       ref.as<uint32_t>(), bvh.tris.as<TriRecord>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(ev1, s));
    HIP_CHECK(hipEventSynchronize(ev1));
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
    bvh.build_ms = ms;
    HIP_CHECK(hipEventDestroy(ev0));
    HIP_CHECK(hipEventDestroy(ev1));
    uint32_t nodes_used = 1;
    if (n > 1)
        HIP_CHECK(hipMemcpy(&nodes_used, counters.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost));
    finalize_bvh(ctx, (int32_t)nodes_used);
}

} // namespace mipt
