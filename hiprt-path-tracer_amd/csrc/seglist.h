// seglist.h -- append-only lists of slot indices written by many waves at once.
//
// A single shared counter serialises every wave's atomic on one cache line;
// with a wave per 64 paths that costs more than the shading itself. Here the
// list is split into SEG_COUNT segments with their own counters, 128 bytes
// apart: wave w appends to segment w % SEG_COUNT (one atomic per wave), and a
// reader maps a dense index to (segment, offset) with a wave-wide prefix sum
// of the SEG_COUNT counts.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mipt {

constexpr uint32_t SEG_COUNT = 64;   // one segment per lane of the reader's scan
constexpr uint32_t SEG_STRIDE = 32;  // uint32 between counters (128 bytes)
constexpr uint32_t SEG_COUNTER_WORDS = SEG_COUNT * SEG_STRIDE;

// Slots a list needs when n threads (64 per wave, consecutive waves in
// consecutive segments) may each append once: every segment receives at most
// ceil(waves / SEG_COUNT) full waves.
__host__ __device__ inline uint32_t seg_capacity(uint32_t n)
{
    const uint32_t waves = (n + 63u) / 64u;
    return (waves + SEG_COUNT - 1u) / SEG_COUNT * 64u;
}

struct SegList {
    uint32_t* slots;    // SEG_COUNT * seg_cap entries
    uint32_t* counts;   // SEG_COUNT counters, SEG_STRIDE apart
    uint32_t seg_cap;
};

#ifdef __HIPCC__
// Appends `value` for every lane with pred; one atomic per wave. Every lane
// of the wave must call it. Returns nothing: readers go through SegMap.
__device__ __forceinline__ void seg_append(const SegList& l, bool pred, uint32_t value)
{
    const uint64_t mask = __ballot(pred);
    if (mask == 0)
        return;
    const uint32_t lane = __lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t seg = wave % SEG_COUNT;
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
    const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    uint32_t base = 0u;
    if (lane == leader)
        base = atomicAdd(l.counts + seg * SEG_STRIDE, (uint32_t)__popcll(mask));
    base = __shfl(base, (int)leader, 64);
    if (pred)
        l.slots[(size_t)seg * l.seg_cap + base + rank] = value;
}

// Dense index -> slot of a segmented list. init() is a wave-wide scan of the
// counts (all 64 lanes must call it); slot() a 6-step binary search over the
// segment starts (all lanes must call it too).
struct SegMap {
    uint32_t excl;  // start of this lane's segment in the dense order
    uint32_t total;
    uint32_t seg_cap;

    __device__ __forceinline__ void init(const uint32_t* counts, uint32_t cap)
    {
        const uint32_t lane = __lane_id();
        const uint32_t c = counts[lane * SEG_STRIDE];
        uint32_t incl = c;
        for (int off = 1; off < 64; off <<= 1)
        {
            const uint32_t t = __shfl_up(incl, off, 64);
            if (lane >= (uint32_t)off)
                incl += t;
        }
        total = __shfl(incl, 63, 64);
        excl = incl - c;
        seg_cap = cap;
    }

    // Position of dense index w (< total) in the slot array.
    __device__ __forceinline__ size_t slot(uint32_t w) const
    {
        int lo = 0, hi = 63;
        for (int step = 0; step < 6; step++)
        {
            const int mid = (lo + hi + 1) >> 1;
            const uint32_t e = __shfl(excl, mid, 64);
            if (e <= w)
                lo = mid;
            else
                hi = mid - 1;
        }
        const uint32_t start = __shfl(excl, lo, 64);
        return (size_t)lo * seg_cap + (w - start);
    }
};
#endif

} // namespace mipt
