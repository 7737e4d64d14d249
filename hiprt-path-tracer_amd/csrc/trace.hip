// trace.hip -- closest-hit and any-hit BVH traversal for gfx950.
//
// Replaces hiprtGeomTraversalClosest / hiprtGeomTraversalAnyHit as used by
// trace_main_path_ray, evaluate_shadow_ray_occluded and
// evaluate_bsdf_light_sample_ray (src/Device/includes/Intersect.h:200-324,
// :544-643). CDNA4 has no ray-tracing units, so this is a plain stackful
// traversal:
//   * one 64-byte BvhNode fetch per step tests both child boxes;
//   * leaf children are intersected immediately (Moller-Trumbore, exact
//     reference arithmetic), interior children are visited near-first and the
//     far one pushed;
//   * the per-lane stack lives in LDS, laid out [depth][lane] so the 64 lanes
//     of a wave touch 64 consecutive dwords (no bank conflicts); entries
//     deeper than the LDS stack spill to a per-thread global array;
//   * a persistent grid sized from occupancy; each wave pulls batches of 64
//     rays from a global counter so that long rays do not stall a whole block;
//   * optionally the rays to trace come from a segmented queue of slot
//     indices (seglist.h) whose length is read on the device (the path passes
//     enqueue only live rays); results go to the rays' own slots.
//
// Results are exact: the box test is widened (boxes padded at build time and
// the slab interval relatively widened by 2^-16) so it never rejects a node
// holding the closest Moller-Trumbore hit, and ties in t go to the smaller
// primitive index. Hit IDs and t are therefore bit-identical to a brute-force
// loop over all triangles with the same arithmetic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <utility>

#include "internal.h"
#include "pt_math.h"
#include "seglist.h"

namespace mipt {
namespace {

enum Mode { CLOSEST = 0, ANY = 1, CLOSEST_COUNTED = 2 };

constexpr float WIDEN_LO = 1.0f - 1.52587890625e-05f; // 1 - 2^-16
constexpr float WIDEN_HI = 1.0f + 1.52587890625e-05f;

__device__ __forceinline__ float safe_inv(float d)
{
    const float tiny = 1.0e-20f;
    return 1.0f / (fabsf(d) > tiny ? d : copysignf(tiny, d));
}

// Slab test of one box; returns the widened entry distance or +inf on a miss.
__device__ __forceinline__ float box_enter(float lox, float loy, float loz, float hix, float hiy, float hiz,
                                           pt::f3 inv, pt::f3 oinv, float tfar_limit)
{
    float tx0 = __builtin_fmaf(lox, inv.x, -oinv.x), tx1 = __builtin_fmaf(hix, inv.x, -oinv.x);
    float ty0 = __builtin_fmaf(loy, inv.y, -oinv.y), ty1 = __builtin_fmaf(hiy, inv.y, -oinv.y);
    float tz0 = __builtin_fmaf(loz, inv.z, -oinv.z), tz1 = __builtin_fmaf(hiz, inv.z, -oinv.z);
    float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tfar_limit));
    return (tn * WIDEN_LO <= tf * WIDEN_HI) ? tn : INFINITY;
}

__global__ void k_reset_counter(uint32_t* c) { c[0] = 0u; }

// Result of one ray's traversal.
struct RayHit {
    int32_t prim;
    float t, u, v;
    uint32_t nodes, tris;
};

// Traverses one ray (the 32-byte record ra, rb) from the root. `any` stops at
// the first accepted hit (any-hit rays are shadow rays: maxT is the light
// distance). The LDS stack is [depth][thread] over the block; entries past
// STACK go to the thread's spill column.
template <int BLOCK, int STACK, bool COUNT>
__device__ __forceinline__ RayHit traverse(const BvhNode* __restrict__ nodes, const TriRecord* __restrict__ tris,
                                           int32_t root, float4 ra, float4 rb, bool any, int32_t* stack,
                                           int32_t* __restrict__ spill, uint32_t spill_stride, uint32_t gtid,
                                           uint32_t max_steps)
{
    const pt::f3 o = pt::mk(ra.x, ra.y, ra.z);
    const pt::f3 d = pt::mk(rb.x, rb.y, rb.z);
    const int32_t last_prim = __float_as_int(ra.w);
    const pt::f3 inv = pt::mk(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    const pt::f3 oinv = pt::mk(o.x * inv.x, o.y * inv.y, o.z * inv.z);

    float best_t = rb.w;
    int32_t best_prim = -1;
    float best_u = 0.0f, best_v = 0.0f;
    uint32_t n_nodes = 0, n_tris = 0;
    bool done = false;

    int32_t node = root;
    int sp = 0;
    // A tree traversal visits each interior node at most once; the cap
    // only guards against a malformed node array.
    for (uint32_t step = 0; !done && step < max_steps; step++)
    {
        const BvhNode* np = nodes + node;
        const float4 a = np->a, b = np->b, c = np->c;
        const int4 refs = np->d;
        if (COUNT)
            n_nodes++;
        const float t0 = box_enter(a.x, a.y, a.z, a.w, b.x, b.y, inv, oinv, best_t);
        const float t1 = box_enter(b.z, b.w, c.x, c.y, c.z, c.w, inv, oinv, best_t);
        const int32_t ch[2] = {refs.x, refs.y};
        const float tc[2] = {t0, t1};

        // Leaf children first.
        for (int s = 0; s < 2 && !done; s++)
        {
            if (ch[s] >= 0 || tc[s] == INFINITY)
                continue;
            const int32_t first = leaf_first(ch[s]);
            const int32_t cnt = leaf_count(ch[s]);
            for (int32_t k = 0; k < cnt; k++)
            {
                const TriRecord* tr = tris + first + k;
                const float4 v0 = tr->v0, e1 = tr->e1, e2 = tr->e2;
                const int32_t prim = __float_as_int(v0.w);
                if (COUNT)
                    n_tris++;
                float t, u, v;
                if (!pt::mt_intersect(pt::mk(v0.x, v0.y, v0.z), pt::mk(e1.x, e1.y, e1.z), pt::mk(e2.x, e2.y, e2.z),
                                      o, d, t, u, v))
                    continue;
                if (prim == last_prim)
                    continue;
                if (t < best_t || (t == best_t && prim < best_prim))
                {
                    best_t = t;
                    best_prim = prim;
                    best_u = u;
                    best_v = v;
                    if (any)
                    {
                        done = true;
                        break;
                    }
                }
            }
        }
        if (done)
            break;

        // Interior children: near first, far pushed.
        const bool h0 = ch[0] >= 0 && t0 != INFINITY;
        const bool h1 = ch[1] >= 0 && t1 != INFINITY;
        int32_t next = -1;
        if (h0 && h1)
        {
            const bool first0 = t0 <= t1;
            next = first0 ? ch[0] : ch[1];
            const int32_t far = first0 ? ch[1] : ch[0];
            if (sp < STACK)
                stack[sp * BLOCK + threadIdx.x] = far;
            else
                spill[(size_t)(sp - STACK) * spill_stride + gtid] = far;
            sp++;
        }
        else if (h0)
            next = ch[0];
        else if (h1)
            next = ch[1];

        if (next < 0)
        {
            if (sp == 0)
                break;
            sp--;
            next = sp < STACK ? stack[sp * BLOCK + threadIdx.x] : spill[(size_t)(sp - STACK) * spill_stride + gtid];
        }
        node = next;
    }
    return RayHit{best_prim, best_t, best_u, best_v, n_nodes, n_tris};
}

__device__ __forceinline__ int4 hit_record(const RayHit& h)
{
    return h.prim >= 0 ? make_int4(h.prim, __float_as_int(h.t), __float_as_int(h.u), __float_as_int(h.v))
                       : make_int4(-1, __float_as_int(-1.0f), 0, 0);
}

// Next tile of a wave: `tile` consecutive batches of 64 rays (the queue
// tile). Each wave's first tile is its own by index; later tiles come from
// the work counter, past the statically assigned ones. A launch on fewer
// rays than the grid's lanes then touches the counter only from the waves
// that had work: a same-address atomic from every wave costs tens of
// microseconds per launch. Larger tiles take fewer atomics and balance
// the waves' uneven ray lengths less finely.
struct BatchCursor {
    uint32_t next;
    bool first;

    __device__ __forceinline__ void init(int block, uint32_t tile)
    {
        next = (blockIdx.x * ((uint32_t)block / 64u) + threadIdx.x / 64u) * 64u * tile;
        first = true;
    }

    __device__ __forceinline__ uint32_t take(uint32_t* work_counter, int block, uint32_t tile)
    {
        uint32_t base = next;
        if (!first)
        {
            const uint32_t waves = gridDim.x * ((uint32_t)block / 64u);
            if ((threadIdx.x & 63u) == 0)
                base = atomicAdd(work_counter, 64u * tile) + waves * 64u * tile;
            base = __shfl(base, 0, 64);
        }
        first = false;
        return base;
    }
};

template <int BLOCK, int STACK, int MODE>
__global__ __launch_bounds__(BLOCK) void k_trace(const BvhNode* __restrict__ nodes, const TriRecord* __restrict__ tris,
                                                 int32_t root, const float4* __restrict__ rays, int4* __restrict__ hits,
                                                 uint8_t* __restrict__ occluded, uint2* __restrict__ counts, uint32_t n,
                                                 int32_t* __restrict__ spill, uint32_t spill_stride,
                                                 uint32_t* __restrict__ work_counter, uint32_t max_steps,
                                                 const uint32_t* __restrict__ queue,
                                                 const uint32_t* __restrict__ queue_counts, uint32_t queue_seg_cap,
                                                 uint32_t tile)
{
    __shared__ int32_t stack[STACK * BLOCK];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    SegMap map{};
    if (queue)
    {
        map.init(queue_counts, queue_seg_cap);
        n = min(n, map.total);
    }
    BatchCursor cur;
    cur.init(BLOCK, tile);
    for (;;)
    {
        const uint32_t first = cur.take(work_counter, BLOCK, tile);
        if (first >= n)
            break;
        for (uint32_t k = 0; k < tile; k++)
        {
            const uint32_t base = first + 64u * k;
            if (base >= n)
                break;
            const uint32_t w = base + lane;
            uint32_t r = w;
            if (queue)
                r = queue[map.slot(w < n ? w : n - 1u)]; // every lane takes part in the search
            if (w >= n)
                continue;
            const RayHit h = traverse<BLOCK, STACK, MODE == CLOSEST_COUNTED>(
                nodes, tris, root, rays[2 * (size_t)r], rays[2 * (size_t)r + 1], MODE == ANY, stack, spill,
                spill_stride, gtid, max_steps);
            if (MODE == ANY)
                occluded[r] = h.prim >= 0 ? 1 : 0;
            else
            {
                hits[r] = hit_record(h);
                if (MODE == CLOSEST_COUNTED)
                    counts[r] = make_uint2(h.nodes, h.tris);
            }
        }
    }
}

// The rays of up to three segmented queues in one launch: dense index w runs
// over queue 0, then 1, then 2, each range rounded up to whole waves so that
// a wave traces one queue; a queue with `occluded` set is traced any-hit, the
// others closest-hit. One launch instead of three keeps the grid busy when
// the queues are short (a launch costs about the traversal latency of its
// longest ray, whatever its length).
struct QueueJob {
    const float4* rays;
    int4* hits;
    uint8_t* occluded;
    const uint32_t* slots;
    const uint32_t* counts;
    uint32_t seg_cap;
};

struct QueueJobs {
    QueueJob job[3];
};

template <int BLOCK, int STACK>
__global__ __launch_bounds__(BLOCK) void k_trace_queues(const BvhNode* __restrict__ nodes,
                                                        const TriRecord* __restrict__ tris, int32_t root, QueueJobs q,
                                                        int32_t* __restrict__ spill, uint32_t spill_stride,
                                                        uint32_t* __restrict__ work_counter, uint32_t max_steps,
                                                        uint32_t tile)
{
    __shared__ int32_t stack[STACK * BLOCK];
    // Segment starts of the three queues (the same for every wave): read
    // from LDS by a per-lane binary search, so no registers hold them while
    // the rays are traversed.
    __shared__ uint32_t seg_start[3][SEG_COUNT];
    __shared__ uint32_t queue_end[3];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    if (threadIdx.x < 64u)
    {
        uint32_t total = 0;
#pragma unroll
        for (int k = 0; k < 3; k++)
        {
            SegMap m{};
            m.excl = 0u;
            m.total = 0u;
            if (q.job[k].slots)
                m.init(q.job[k].counts, q.job[k].seg_cap);
            seg_start[k][lane] = m.excl;
            total += m.total;
            if (lane == 0)
                queue_end[k] = total;
        }
    }
    __syncthreads();
    // each queue's range starts on a whole wave
    const uint32_t n0 = queue_end[0], n1 = queue_end[1] - n0, n2 = queue_end[2] - queue_end[1];
    const uint32_t start1 = (n0 + 63u) & ~63u;
    const uint32_t start2 = start1 + ((n1 + 63u) & ~63u);
    const uint32_t total = start2 + n2;
    BatchCursor cur;
    cur.init(BLOCK, tile);
    for (;;)
    {
        const uint32_t first = cur.take(work_counter, BLOCK, tile);
        if (first >= total)
            break;
        for (uint32_t kt = 0; kt < tile; kt++)
        {
            const uint32_t base = first + 64u * kt;
            if (base >= total)
                break;
            // the wave's queue (uniform over the wave)
            const int k = base < start1 ? 0 : (base < start2 ? 1 : 2);
            const uint32_t local = base + lane - (k == 0 ? 0u : (k == 1 ? start1 : start2));
            const uint32_t count = k == 0 ? n0 : (k == 1 ? n1 : n2);
            if (local >= count)
                continue;
            // fields picked by selects: indexing the argument struct by a
            // run-time value would copy it to scratch memory
            const float4* rays = k == 0 ? q.job[0].rays : (k == 1 ? q.job[1].rays : q.job[2].rays);
            const uint32_t* slots = k == 0 ? q.job[0].slots : (k == 1 ? q.job[1].slots : q.job[2].slots);
            const uint32_t seg_cap = k == 0 ? q.job[0].seg_cap : (k == 1 ? q.job[1].seg_cap : q.job[2].seg_cap);
            // last segment whose start is <= local
            int lo = 0, hi = (int)SEG_COUNT - 1;
            for (int step = 0; step < 6; step++)
            {
                const int mid = (lo + hi + 1) >> 1;
                if (seg_start[k][mid] <= local)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const uint32_t r = slots[(size_t)lo * seg_cap + (local - seg_start[k][lo])];
            const float4 ra = rays[2 * (size_t)r], rb = rays[2 * (size_t)r + 1];
            uint8_t* occluded = k == 0 ? q.job[0].occluded : (k == 1 ? q.job[1].occluded : q.job[2].occluded);
            if (occluded)
            {
                const RayHit h = traverse<BLOCK, STACK, false>(nodes, tris, root, ra, rb, true, stack, spill,
                                                              spill_stride, gtid, max_steps);
                occluded[r] = h.prim >= 0 ? 1 : 0;
            }
            else
            {
                const RayHit h = traverse<BLOCK, STACK, false>(nodes, tris, root, ra, rb, false, stack, spill,
                                                              spill_stride, gtid, max_steps);
                int4* hits = k == 0 ? q.job[0].hits : (k == 1 ? q.job[1].hits : q.job[2].hits);
                hits[r] = hit_record(h);
            }
        }
    }
}

__global__ void k_all_miss(int4* __restrict__ hits, uint8_t* __restrict__ occluded, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    if (hits)
        hits[i] = make_int4(-1, __float_as_int(-1.0f), 0, 0);
    if (occluded)
        occluded[i] = 0;
}

using TraceFn = void (*)(const BvhNode*, const TriRecord*, int32_t, const float4*, int4*, uint8_t*, uint2*,
                         uint32_t, int32_t*, uint32_t, uint32_t*, uint32_t, const uint32_t*, const uint32_t*, uint32_t,
                         uint32_t);

template <int BLOCK, int STACK>
TraceFn pick_mode(int mode)
{
    switch (mode)
    {
    case CLOSEST: return k_trace<BLOCK, STACK, CLOSEST>;
    case ANY: return k_trace<BLOCK, STACK, ANY>;
    default: return k_trace<BLOCK, STACK, CLOSEST_COUNTED>;
    }
}

template <int BLOCK>
TraceFn pick_stack(int stack, int mode)
{
    switch (stack)
    {
    case 8: return pick_mode<BLOCK, 8>(mode);
    case 16: return pick_mode<BLOCK, 16>(mode);
    case 24: return pick_mode<BLOCK, 24>(mode);
    default: return pick_mode<BLOCK, 32>(mode);
    }
}

TraceFn pick(int block, int stack, int mode)
{
    switch (block)
    {
    case 64: return pick_stack<64>(stack, mode);
    case 128: return pick_stack<128>(stack, mode);
    default: return pick_stack<256>(stack, mode);
    }
}

struct Launch {
    TraceFn fn;
    unsigned grid;
};

using QueuesFn = void (*)(const BvhNode*, const TriRecord*, int32_t, QueueJobs, int32_t*, uint32_t, uint32_t*, uint32_t,
                          uint32_t);

template <int BLOCK>
QueuesFn pick_queues_stack(int stack)
{
    switch (stack)
    {
    case 8: return k_trace_queues<BLOCK, 8>;
    case 16: return k_trace_queues<BLOCK, 16>;
    case 24: return k_trace_queues<BLOCK, 24>;
    default: return k_trace_queues<BLOCK, 32>;
    }
}

QueuesFn pick_queues(int block, int stack)
{
    switch (block)
    {
    case 64: return pick_queues_stack<64>(stack);
    case 128: return pick_queues_stack<128>(stack);
    default: return pick_queues_stack<256>(stack);
    }
}

// Persistent grid: the blocks that fit on the device at once, at most enough
// for n rays.
unsigned persistent_grid(Ctx& ctx, const void* fn, uint32_t n)
{
    const int block = ctx.trace_cfg.block_size;
    // Blocks per CU of each kernel on each device, computed once; contexts on
    // several host threads share the table.
    static std::map<std::pair<const void*, int>, int> occupancy;
    static std::mutex occupancy_lock;
    const auto key = std::make_pair(fn, ctx.device);
    int per_cu = 0;
    {
        std::lock_guard<std::mutex> guard(occupancy_lock);
        auto it = occupancy.find(key);
        if (it != occupancy.end())
            per_cu = it->second;
    }
    if (per_cu == 0)
    {
        HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, 0));
        if (per_cu < 1)
            per_cu = 1;
        std::lock_guard<std::mutex> guard(occupancy_lock);
        occupancy[key] = per_cu;
    }
    const uint64_t max_grid = (uint64_t)per_cu * (uint64_t)ctx.cu_count;
    const uint64_t need = ((uint64_t)n + block - 1) / block;
    const uint64_t g = need < max_grid ? need : max_grid;
    return (unsigned)(g ? g : 1);
}

Launch plan(Ctx& ctx, int mode, uint32_t n)
{
    Launch l;
    l.fn = pick(ctx.trace_cfg.block_size, ctx.trace_cfg.lds_stack, mode);
    l.grid = persistent_grid(ctx, reinterpret_cast<const void*>(l.fn), n);
    return l;
}

// n: rays (or, with a queue, the queue's capacity: slots [0, n) of rays).
// Work counter and spill columns of a persistent grid of `threads` threads.
struct Scratch {
    uint32_t* counter;
    int32_t* spill;
};

Scratch trace_scratch(Ctx& ctx, uint32_t threads)
{
    // Spill entries: the stack never holds more than max_depth entries. The
    // columns are sized for every thread the device can hold (32 waves per
    // CU), not for this launch's grid, so that launches of different grids
    // on one stream never reallocate the buffer under queued kernels.
    const int spill_depth = ctx.bvh.max_depth - ctx.trace_cfg.lds_stack;
    const size_t spill_off = 256; // work counter lives at the front
    const size_t columns = std::max<size_t>(threads, (size_t)ctx.cu_count * 32u * 64u);
    size_t spill_bytes = spill_off;
    if (spill_depth > 0)
        spill_bytes = spill_off + sizeof(int32_t) * (size_t)spill_depth * columns;
    ctx.spill.reserve(spill_bytes);
    return Scratch{ctx.spill.as<uint32_t>(),
                   reinterpret_cast<int32_t*>(reinterpret_cast<char*>(ctx.spill.ptr) + spill_off)};
}

bool trace_ready(Ctx& ctx)
{
    if (!ctx.bvh.valid)
        fail(MIPT_ERR_NO_BVH, "no BVH: call mipt_build_bvh first");
    return ctx.bvh.triangle_count != 0;
}

void launch(Ctx& ctx, int mode, const mipt_ray* rays, mipt_hit* hits, uint8_t* occluded, uint32_t* counts, uint32_t n,
            const SegList* queue, hipStream_t stream)
{
    if (n == 0)
        return;
    if (!trace_ready(ctx))
    {
        hipLaunchKernelGGL(k_all_miss, dim3((n + 255) / 256), dim3(256), 0, stream, reinterpret_cast<int4*>(hits),
                           occluded, n);
        HIP_CHECK(hipGetLastError());
        return;
    }
    const Launch l = plan(ctx, mode, n);
    const int block = ctx.trace_cfg.block_size;
    const uint32_t threads = l.grid * (uint32_t)block;
    const Scratch sc = trace_scratch(ctx, threads);
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, stream, sc.counter);
    hipLaunchKernelGGL(l.fn, dim3(l.grid), dim3(block), 0, stream, ctx.bvh.nodes.as<BvhNode>(),
                       ctx.bvh.tris.as<TriRecord>(), ctx.bvh.root, reinterpret_cast<const float4*>(rays),
                       reinterpret_cast<int4*>(hits), occluded, reinterpret_cast<uint2*>(counts), n, sc.spill, threads,
                       sc.counter, (uint32_t)ctx.bvh.node_count + 1u, queue ? queue->slots : nullptr,
                       queue ? queue->counts : nullptr, queue ? queue->seg_cap : 0u,
                       (uint32_t)ctx.trace_cfg.queue_tile);
    HIP_CHECK(hipGetLastError());
}

// ---- closest hit without allocation ---------------------------------
// Work counter and spill columns of mipt_trace_closest_noalloc, in a
// file-scope device array: each device holds its own copy from module load
// until the process ends. Host code reaches it only through the per-device
// address from hipGetSymbolAddress (the name itself is a host stand-in).
constexpr size_t NOALLOC_COUNTER_BYTES = 256;
static_assert(MIPT_TRACE_NOALLOC_SCRATCH_BYTES > NOALLOC_COUNTER_BYTES, "scratch too small");
static __device__ __align__(256) uint8_t scratch[MIPT_TRACE_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex noalloc_lock;
void* noalloc_ptr[MAX_DEVICES] = {};
hipError_t noalloc_err[MAX_DEVICES] = {};
bool noalloc_done[MAX_DEVICES] = {};

} // namespace

void noalloc_scratch_init(int device)
{
    if (device < 0 || device >= MAX_DEVICES)
        return;
    std::lock_guard<std::mutex> guard(noalloc_lock);
    if (noalloc_done[device])
        return;
    // the device is current (mipt_create set it)
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    noalloc_err[device] = e;
    noalloc_ptr[device] = e == hipSuccess ? p : nullptr;
    noalloc_done[device] = true;
}

void trace_closest_noalloc(Ctx& ctx, const mipt_ray* rays, mipt_hit* hits, uint32_t n, hipStream_t stream)
{
    if (n == 0)
        return;
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> guard(noalloc_lock);
        if (ctx.device < 0 || ctx.device >= MAX_DEVICES || !noalloc_done[ctx.device])
            fail(MIPT_ERR_HIP, "mipt_trace_closest_noalloc: no scratch address for this device");
        if (noalloc_err[ctx.device] != hipSuccess)
            fail(MIPT_ERR_HIP, std::string("mipt_trace_closest_noalloc: hipGetSymbolAddress(scratch): ") +
                                   hipGetErrorString(noalloc_err[ctx.device]));
        base = noalloc_ptr[ctx.device];
    }
    if (!trace_ready(ctx))
    {
        hipLaunchKernelGGL(k_all_miss, dim3((n + 255) / 256), dim3(256), 0, stream, reinterpret_cast<int4*>(hits),
                           nullptr, n);
        HIP_CHECK(hipGetLastError());
        return;
    }
    const Launch l = plan(ctx, CLOSEST, n);
    const int block = ctx.trace_cfg.block_size;
    // Spill columns of the persistent grid: when the tree is deep enough
    // that the whole grid's columns do not fit, the grid shrinks to the
    // blocks whose columns do (every wave takes rays from the work counter,
    // so fewer blocks trace the same rays).
    const int spill_depth = std::max(0, ctx.bvh.max_depth - ctx.trace_cfg.lds_stack);
    const size_t room = MIPT_TRACE_NOALLOC_SCRATCH_BYTES - NOALLOC_COUNTER_BYTES;
    unsigned grid = l.grid;
    if (spill_depth > 0)
    {
        const size_t per_block = sizeof(int32_t) * (size_t)spill_depth * (size_t)block;
        const size_t fit = room / per_block;
        if (fit == 0)
            fail(MIPT_ERR_INVALID_ARGUMENT,
                 "mipt_trace_closest_noalloc: a BVH of depth " + std::to_string(ctx.bvh.max_depth) +
                     " needs more than the " + std::to_string(MIPT_TRACE_NOALLOC_SCRATCH_BYTES) +
                     "-byte scratch for one block's spill stack");
        grid = (unsigned)std::min<size_t>(grid, fit);
    }
    const uint32_t threads = grid * (uint32_t)block;
    uint32_t* counter = reinterpret_cast<uint32_t*>(base);
    int32_t* spill = reinterpret_cast<int32_t*>(reinterpret_cast<char*>(base) + NOALLOC_COUNTER_BYTES);
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, stream, counter);
    hipLaunchKernelGGL(l.fn, dim3(grid), dim3(block), 0, stream, ctx.bvh.nodes.as<BvhNode>(),
                       ctx.bvh.tris.as<TriRecord>(), ctx.bvh.root, reinterpret_cast<const float4*>(rays),
                       reinterpret_cast<int4*>(hits), nullptr, nullptr, n, spill, threads, counter,
                       (uint32_t)ctx.bvh.node_count + 1u, nullptr, nullptr, 0u, (uint32_t)ctx.trace_cfg.queue_tile);
    HIP_CHECK(hipGetLastError());
}

void trace_reserve(Ctx& ctx)
{
    if (ctx.bvh.valid && ctx.bvh.triangle_count > 0)
        (void)trace_scratch(ctx, 0u);
}

void autotune_trace(Ctx& ctx, const mipt_ray* rays, uint32_t n, int32_t repeats, hipStream_t stream,
                    mipt_trace_tuning* out, double* table)
{
    static const int blocks[3] = {64, 128, 256};
    static const int stacks[4] = {8, 16, 24, 32};
    static const int tiles[3] = {1, 2, 4};
    if (!trace_ready(ctx))
        fail(MIPT_ERR_INVALID_ARGUMENT, "mipt_autotune_trace: the scene has no triangles");
    const TraceConfig keep = ctx.trace_cfg;
    ctx.autotune_hits.reserve(sizeof(mipt_hit) * (size_t)n);
    mipt_hit* hits = ctx.autotune_hits.as<mipt_hit>();
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    double best = 0.0, at_default = 0.0;
    TraceConfig best_cfg = keep;
    int tried = 0;
    try
    {
        for (int bi = 0; bi < 3; bi++)
            for (int si = 0; si < 4; si++)
                for (int ti = 0; ti < 3; ti++)
                {
                    ctx.trace_cfg.block_size = blocks[bi];
                    ctx.trace_cfg.lds_stack = stacks[si];
                    ctx.trace_cfg.queue_tile = tiles[ti];
                    // one untimed launch: occupancy query, spill columns
                    launch(ctx, CLOSEST, rays, hits, nullptr, nullptr, n, nullptr, stream);
                    HIP_CHECK(hipEventRecord(e0, stream));
                    for (int r = 0; r < repeats; r++)
                        launch(ctx, CLOSEST, rays, hits, nullptr, nullptr, n, nullptr, stream);
                    HIP_CHECK(hipEventRecord(e1, stream));
                    HIP_CHECK(hipEventSynchronize(e1));
                    float ms = 0.0f;
                    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                    const double mean = (double)ms / repeats;
                    if (table)
                        table[(bi * 4 + si) * 3 + ti] = mean;
                    if (blocks[bi] == 128 && stacks[si] == 16 && tiles[ti] == 1)
                        at_default = mean;
                    if (tried == 0 || mean < best)
                    {
                        best = mean;
                        best_cfg = ctx.trace_cfg;
                    }
                    tried++;
                }
    }
    catch (...)
    {
        ctx.trace_cfg = keep;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
    ctx.trace_cfg = best_cfg;
    if (out)
    {
        out->block_size = best_cfg.block_size;
        out->lds_stack = best_cfg.lds_stack;
        out->queue_tile = best_cfg.queue_tile;
        out->configurations = tried;
        out->best_ms = best;
        out->default_ms = at_default;
    }
}

void trace_closest(Ctx& ctx, const mipt_ray* rays, mipt_hit* hits, uint32_t* counts, uint32_t n, hipStream_t stream)
{
    launch(ctx, counts ? CLOSEST_COUNTED : CLOSEST, rays, hits, nullptr, counts, n, nullptr, stream);
}

void trace_any(Ctx& ctx, const mipt_ray* rays, uint8_t* occluded, uint32_t n, hipStream_t stream)
{
    launch(ctx, ANY, rays, nullptr, occluded, nullptr, n, nullptr, stream);
}

void trace_closest_queue(Ctx& ctx, const mipt_ray* rays, mipt_hit* hits, const SegList& queue, uint32_t capacity,
                         hipStream_t stream)
{
    launch(ctx, CLOSEST, rays, hits, nullptr, nullptr, capacity, &queue, stream);
}

void trace_any_queue(Ctx& ctx, const mipt_ray* rays, uint8_t* occluded, const SegList& queue, uint32_t capacity,
                     hipStream_t stream)
{
    launch(ctx, ANY, rays, nullptr, occluded, nullptr, capacity, &queue, stream);
}

void trace_queues(Ctx& ctx, const TraceQueue* q, int count, uint32_t capacity, hipStream_t stream)
{
    if (capacity == 0 || count == 0)
        return;
    if (count > 3)
        fail(MIPT_ERR_INVALID_ARGUMENT, "trace_queues: at most 3 queues");
    if (!trace_ready(ctx))
    {
        for (int k = 0; k < count; k++)
            hipLaunchKernelGGL(k_all_miss, dim3((capacity + 255) / 256), dim3(256), 0, stream,
                               reinterpret_cast<int4*>(q[k].hits), q[k].occluded, capacity);
        HIP_CHECK(hipGetLastError());
        return;
    }
    QueueJobs jobs{};
    for (int k = 0; k < count; k++)
    {
        QueueJob& j = jobs.job[k];
        j.rays = reinterpret_cast<const float4*>(q[k].rays);
        j.hits = reinterpret_cast<int4*>(q[k].hits);
        j.occluded = q[k].occluded;
        j.slots = q[k].queue->slots;
        j.counts = q[k].queue->counts;
        j.seg_cap = q[k].queue->seg_cap;
    }
    const int block = ctx.trace_cfg.block_size;
    const QueuesFn fn = pick_queues(block, ctx.trace_cfg.lds_stack);
    const unsigned grid = persistent_grid(ctx, reinterpret_cast<const void*>(fn), (uint32_t)std::min<uint64_t>(
                                                                                     (uint64_t)capacity * count,
                                                                                     0xffffffffull));
    const uint32_t threads = grid * (uint32_t)block;
    const Scratch sc = trace_scratch(ctx, threads);
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, stream, sc.counter);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(block), 0, stream, ctx.bvh.nodes.as<BvhNode>(),
                       ctx.bvh.tris.as<TriRecord>(), ctx.bvh.root, jobs, sc.spill, threads, sc.counter,
                       (uint32_t)ctx.bvh.node_count + 1u, (uint32_t)ctx.trace_cfg.queue_tile);
    HIP_CHECK(hipGetLastError());
}

} // namespace mipt
