// internal.h -- context, device buffers and launch helpers of libmipt.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/mipt.h"
#include "layout.h"

namespace mipt {

// Thrown inside the library only; every C entry point catches it and turns it
// into a negative status + message.
struct Error {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg);
void hip_check(hipError_t e, const char* what);

#define HIP_CHECK(expr) ::mipt::hip_check((expr), #expr)

// Owning device allocation (hipMalloc), sized in bytes.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void release();
    // Grows (never shrinks) to at least `n` bytes; contents are not kept.
    void reserve(size_t n);
    template <class T> T* as() const { return reinterpret_cast<T*>(ptr); }
};

struct TraceConfig {
    int block_size = 128;
    int lds_stack = 16;
};

struct Bvh {
    bool valid = false;
    int32_t builder = 0;
    int32_t triangle_count = 0;
    int32_t node_count = 0;
    int32_t leaf_count = 0;
    int32_t max_depth = 0;
    int32_t max_leaf_size = 0;
    int32_t root = 0;             // node index of the root
    double build_ms = 0.0;
    double sah_cost = 0.0;
    DeviceBuffer nodes;           // BvhNode[node_count]
    DeviceBuffer tris;            // TriRecord[triangle_count] in leaf order
};

struct Ctx {
    int device = 0;
    int cu_count = 0;
    std::string last_error;

    // Scene (device copies of the upload arrays).
    int32_t triangle_count = 0;
    int32_t vertex_count = 0;
    int32_t material_count = 0;
    int32_t emissive_count = 0;
    DeviceBuffer vertices;        // float[3V]
    DeviceBuffer indices;         // int[3T]
    DeviceBuffer has_normals;     // uint8[V]
    DeviceBuffer vertex_normals;  // float[3V]
    DeviceBuffer material_indices;// int[T]
    DeviceBuffer emissive_prims;  // int[E]
    std::vector<mipt_material> materials_host;
    bool scene_set = false;

    Bvh bvh;
    TraceConfig trace_cfg;

    // Traversal spill stack (entries beyond the LDS stack), sized on demand.
    DeviceBuffer spill;
};

// bvh_build.hip
void build_lbvh(Ctx& ctx);

// trace.hip
void trace_closest(Ctx& ctx, const mipt_ray* rays, mipt_hit* hits, uint32_t* counts, uint32_t n,
                   hipStream_t stream);
void trace_any(Ctx& ctx, const mipt_ray* rays, uint8_t* occluded, uint32_t n, hipStream_t stream);

// camera.hip
void camera_rays(Ctx& ctx, const mipt_camera& cam, int32_t w, int32_t h, uint32_t sample_number,
                 uint32_t random_number, int32_t freeze, mipt_ray* rays, uint32_t* rng,
                 hipStream_t stream);
void bounce_rays(Ctx& ctx, const mipt_ray* primary, const mipt_hit* hits, uint32_t* rng, mipt_ray* out,
                 uint32_t n, hipStream_t stream);

} // namespace mipt
