// internal.h -- context, device buffers and launch helpers of libmipt.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/mipt.h"
#include "layout.h"
#include "seglist.h"

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

// Boundary-skip rounds of a bounce at most (nested dielectrics: each skip adds
// a stack entry without a pop, so at most 2 follow one another on a 4-entry
// stack whose entry 0 is the air), and the kernels that append to a bounce's
// list of live paths: its resolve kernel and the skip rounds.
constexpr int32_t MAX_SKIP_ROUNDS = 2;
constexpr uint32_t LIVE_APPENDERS = 1u + (uint32_t)MAX_SKIP_ROUNDS;

// Material-sorted shading order: 64 sort classes (32 lobe classes x the
// path's volume state stored or not), each counted on 16 sub-counters
// (a block b adds to sub-counter b % 16) to spread the atomics.
constexpr uint32_t SHADE_SORT_CLASSES = 64;
constexpr uint32_t SHADE_SORT_SUBS = 16;
constexpr uint32_t SHADE_SORT_COUNTERS = SHADE_SORT_CLASSES * SHADE_SORT_SUBS;

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
    int queue_tile = 1; // rays a wave takes per work-counter atomic, in units of 64
};

// Per-sample values the pass kernels read from device memory when they run
// as a replayed hipGraph (written before each replay by a one-thread kernel);
// eager launches pass them by value instead.
struct FrameParams {
    uint32_t sample_number;
    uint32_t rn_camera;   // host random number of the camera pass
    uint32_t rn_path;     // of the path-tracing pass
    float consistent_tau; // consistent_tau(tau_0, sample_number)
    float consistent_tau_refract;
};

// One camera view of a batch (device copy): the HIPRTCamera matrices the
// ray generation reads, its resolution and its first pixel in the batch.
struct ViewDev {
    float iv[4][4];
    float ip[4][4];
    float position[3]; // the camera position (the first vertex's view direction)
    int32_t jitter;
    int32_t width, height;
    uint32_t start;
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

// Per-resolution render buffers and settings of the path-tracing passes.
struct RenderState {
    bool configured = false;
    bool camera_set = false;
    mipt_render_settings settings{};
    mipt_camera camera{};
    // the camera of the last camera-ray pass, and the one before it (the
    // reprojection of ReSTIR DI temporal reuse when not accumulating)
    mipt_camera frame_camera{}, prev_camera{};
    bool frame_camera_valid = false;
    uint32_t frame_count = 0; // mipt_render_sample calls since the buffers were allocated (not accumulating)
    uint32_t pixels = 0;
    // scene-derived
    DeviceBuffer materials;     // pt::Material[material_count]
    DeviceBuffer mat_class;     // int32[material_count]: lobe class (0..30) of each material
    int32_t lobe_classes = 0;   // distinct lobe classes in the scene
    int32_t shade_order = MIPT_SHADE_ORDER_AUTO;
    DeviceBuffer shade_sorted, sort_counts; // live paths in shading order; class counters and cursors
    // batched camera views (mipt_set_views): the pixels of every view back to
    // back (settings.width = total, height = 1)
    std::vector<mipt_view> views;
    DeviceBuffer view_table;    // ViewDev[views]
    // graph replay of a plain sample (mipt_set_graph_capture)
    bool graph_enabled = false;
    bool graph_dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t capture_stream = nullptr;
    DeviceBuffer frame_params;  // FrameParams
    uint32_t graph_replays = 0;
    DeviceBuffer areas;         // float[T]
    DeviceBuffer lut_conductor; // float[128*128]
    DeviceBuffer lut_glossy, lut_glass, lut_glass_inv, lut_thin_glass, lut_glass_exponent, lut_sheen;
    // tables set with mipt_set_lut (MIPT_LUT_GGX_CONDUCTOR .. MIPT_LUT_GGX_THIN_GLASS), empty: the shipped one
    std::vector<float> lut_override[5];
    bool scene_ready = false;
    bool scene_volumes = false;    // some material transmits light
    bool scene_dispersive = false; // some material is dispersive glass
    // G-buffer (CameraRays output)
    DeviceBuffer gb_prim, gb_pos, gb_gn, gb_sn, gb_mat;
    // path state
    DeviceBuffer throughput, thr_prev, radiance, vtx_pos, vtx_sn, vtx_gn, vtx_view, nee_light, nee_bsdf, rng, flags;
    DeviceBuffer vol_a, vol_b;    // volume state of the paths that keep one (uint4 each)
    DeviceBuffer ris_state;       // 3 float4 per path with RIS of the lights
    // rays of one bounce, and the queues of live slots the traversal takes
    DeviceBuffer shadow_rays, shadow_occ, bsdf_rays, bsdf_hits, cont_rays, cont_hits;
    // envmap sampling: the light sample's and the BSDF sample's pending
    // contributions; their shadow rays take shadow slots n + pix and 2n + pix
    DeviceBuffer env_light, env_bsdf;
    uint32_t shadow_slots = 1;  // shadow ray slots per pixel the buffers hold
    DeviceBuffer q_shadow, q_bsdf, q_cont, q_skip, q_live, q_counts; // q_skip, q_live: two lists back to back
    // framebuffer
    DeviceBuffer fb, fb64, counters;
    // power alias table of the emissive triangles (LSS_BASE_POWER)
    DeviceBuffer alias_probas, alias_alias;
    std::vector<float> alias_probas_host;
    std::vector<int32_t> alias_alias_host;
    float alias_sum = 0.0f;
    // ReSTIR DI (LSS_RESTIR_DI)
    mipt_restir_di_settings restir{};
    DeviceBuffer pg_prim, pg_pos, pg_gn, pg_sn, pg_mat; // previous-frame G-buffer
    DeviceBuffer reservoirs[3];   // initial candidates, spatial output 1, spatial output 2
    int32_t last_output = 1;      // m_last_restir_output_reservoirs (ReSTIRDIRenderPass.cpp:251)
    int32_t restir_shade_buffer = 1; // restir_output_reservoirs of the current frame
    int32_t temporal_output = 0;  // buffer the temporal pass of this frame wrote
    int32_t fused_output = 1;     // buffer the fused pass (or the spatial pass after it) wrote
    bool odd_frame = false;
    // ReSTIR DI pass scratch
    DeviceBuffer rs_state, rs_rays, rs_occ, rs_hits, rs_aux;
    DeviceBuffer presampled;      // presampled lights (light presampling)
    bool presampled_ready = false; // the presampling pass has filled `presampled` since it was allocated
};

// Learned radiance cache state (nrc.hip).
struct Nrc {
    bool configured = false;
    mipt_radiance_cache_settings settings{};
    size_t grid_count = 0;        // levels * table * 2 features
    uint32_t step = 0;            // Adam steps taken
    // row selection of the last gradient batch (the Adam step skips an empty one)
    const uint32_t* last_sel_count = nullptr;
    uint32_t last_sel_stride = 1, last_sel_offset = 0, last_sel_limit = 0;
    DeviceBuffer grid_f16, grid_master, grid_m, grid_v, grid_grad;
    DeviceBuffer w_master, w_m, w_v, w_grad, wt, wk;
    DeviceBuffer x0, xcm, dycm, dx0, loss, count;
    // records of the path pass
    DeviceBuffer queries, query_pix, query_out, counters, train_queries, train_targets, train_pix, train_rad;
    DeviceBuffer query_list, train_list, list_counts; // segmented pixel lists of the query kernel
    // query records in spatial (Morton) order, and the sort's buffers
    DeviceBuffer queries_sorted, query_pix_sorted, sort_keys, sort_keys_out, sort_vals, sort_vals_out, sort_tmp;
    uint32_t train_capacity = 0;  // training records one path pass can hold
};

struct Ctx {
    int device = 0;
    int cu_count = 0;
    std::string last_error;
    float scene_bmin[3] = {0.0f, 0.0f, 0.0f};
    float scene_bmax[3] = {1.0f, 1.0f, 1.0f};
    Nrc nrc;

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
    std::vector<float> host_vertices;   // host copies of the geometry
    std::vector<int32_t> host_indices;
    bool scene_set = false;

    Bvh bvh;
    TraceConfig trace_cfg;

    // Ambient light (mipt_set_world_settings / mipt_set_envmap)
    mipt_world_settings world{};
    int32_t env_width = 0, env_height = 0;
    float env_sum = 0.0f;             // luminance sum of the float texels
    DeviceBuffer env_texels;          // uint32 R9G9B9E5 [env_width * env_height]
    DeviceBuffer env_probas, env_alias;
    std::vector<float> env_probas_host;
    std::vector<int32_t> env_alias_host;

    // Traversal spill stack (entries beyond the LDS stack), sized on demand.
    DeviceBuffer spill;
    DeviceBuffer autotune_hits; // hit records of the autotuner's timed launches

    RenderState render;
};

// materials.cpp: CPUMaterial -> the 8-bit packed form -> unpacked floats
// (pt::Material layout, 64 floats); returns an empty string or the reason a
// material uses a lobe this build does not implement.
std::string unpack_material(const mipt_material& in, float out[64]);
const float* conductor_lut_data(); // 128 x 128, luts.cpp
const float* conductor_correlated_lut_data(); // 128 x 128, height-correlated Smith
const float* glossy_dielectric_lut_data(); // [128][64][128]
const float* glass_lut_data();             // [128][16][256]
const float* glass_inv_lut_data();         // [128][16][256]
const float* thin_glass_lut_data();        // [96][32][32]
const float* glass_exponent_data();        // [2][10][11][4]
const float* sheen_ltc_data();             // [32][32][3]

// pathtrace.hip
// With `frame` (graph capture), the kernels read the sample's values from it.
void camera_rays_pass(Ctx& ctx, uint32_t sample_number, uint32_t random_number, hipStream_t stream,
                      const FrameParams* frame = nullptr);
void path_trace_pass(Ctx& ctx, uint32_t sample_number, uint32_t random_number, hipStream_t stream,
                     bool warmup = false, const FrameParams* frame = nullptr);
float consistent_tau(float tau_0, uint32_t sample_number);
void write_frame_params(FrameParams* frame, const FrameParams& values, hipStream_t stream);

// envmap.hip: the ambient light's envmap and its device view.
void set_envmap(Ctx& ctx, const float* pixels, int32_t width, int32_t height, int32_t channels);
// baker.hip
int32_t bake_default_settings(int32_t table, mipt_bake_settings* s);
void bake_plan(const mipt_bake_settings& s, int32_t& ipk, int32_t& launches, int32_t& normalizer);
void bake_lut(Ctx& ctx, const mipt_bake_settings& s, float* out, mipt_bake_info* info);
bool env_sampling_on(const Ctx& ctx); // ENVMAP ambient light with intensity > 0
void allocate_env_path_buffers(Ctx& ctx);

// lights.cpp: Utils::compute_alias_table (src/Utils/Utils.cpp:63-131) and the
// emissive power list (GPURenderer.cpp:221-295).
void compute_alias_table(const float* weights, int32_t n, float total_sum, float* out_probas, int32_t* out_alias);
void build_power_alias_table(Ctx& ctx);

// restir.hip: the ReSTIR DI passes and their buffers.
bool restir_needs_prev_gbuffer(const RenderState& R);
void allocate_restir_buffers(Ctx& ctx);
void restir_reset(Ctx& ctx, hipStream_t stream);
float4* restir_reservoir_buffer(Ctx& ctx, int32_t which);
void restir_initial_pass(Ctx& ctx, uint32_t sample_number, uint32_t random_number, hipStream_t stream);
void restir_temporal_pass(Ctx& ctx, uint32_t sample_number, uint32_t random_number, hipStream_t stream);
void restir_spatial_pass(Ctx& ctx, int32_t pass_index, uint32_t sample_number, uint32_t random_number,
                         hipStream_t stream);
void restir_presampling_pass(Ctx& ctx, uint32_t sample_number, uint32_t random_number, hipStream_t stream);
void restir_spatiotemporal_pass(Ctx& ctx, uint32_t sample_number, uint32_t random_number, hipStream_t stream);

// bvh_build.hip
void build_lbvh(Ctx& ctx);
void build_sah_binned(Ctx& ctx);

// trace.hip
void trace_closest(Ctx& ctx, const mipt_ray* rays, mipt_hit* hits, uint32_t* counts, uint32_t n,
                   hipStream_t stream);
void trace_any(Ctx& ctx, const mipt_ray* rays, uint8_t* occluded, uint32_t n, hipStream_t stream);
// Sizes the traversal's spill stack for the current tree and launch
// configuration (it is then never reallocated by a launch: graph capture).
void trace_reserve(Ctx& ctx);
// Times closest-hit launches over `rays` for every block size (64, 128,
// 256), LDS stack depth (8, 16, 24, 32) and queue tile (1, 2, 4 batches of
// 64 rays) and keeps the fastest; table (36 entries, [block][stack][tile])
// receives every mean launch time in ms.
void autotune_trace(Ctx& ctx, const mipt_ray* rays, uint32_t n, int32_t repeats, hipStream_t stream,
                    mipt_trace_tuning* out, double* table);
// mipt_trace_closest_noalloc: the closest-hit kernel with its work counter
// and spill stack in the library's static device scratch, whose per-device
// address noalloc_scratch_init looks up (once per device, device current).
void noalloc_scratch_init(int device);
void trace_closest_noalloc(Ctx& ctx, const mipt_ray* rays, mipt_hit* hits, uint32_t n, hipStream_t stream);
// Trace the rays whose slots a segmented queue lists (seglist.h; length read
// on the device, at most capacity rays); results go to the same slots.
void trace_closest_queue(Ctx& ctx, const mipt_ray* rays, mipt_hit* hits, const SegList& queue, uint32_t capacity,
                         hipStream_t stream);
void trace_any_queue(Ctx& ctx, const mipt_ray* rays, uint8_t* occluded, const SegList& queue, uint32_t capacity,
                     hipStream_t stream);
// Up to three queued ray sets in one launch: a set with `occluded` is traced
// any-hit (shadow rays), the others closest-hit into `hits`.
struct TraceQueue {
    const mipt_ray* rays;
    mipt_hit* hits;
    uint8_t* occluded;
    const SegList* queue;
};
void trace_queues(Ctx& ctx, const TraceQueue* queues, int count, uint32_t capacity, hipStream_t stream);

// camera.hip
void camera_rays(Ctx& ctx, const mipt_camera& cam, int32_t w, int32_t h, uint32_t sample_number,
                 uint32_t random_number, int32_t freeze, mipt_ray* rays, uint32_t* rng,
                 hipStream_t stream, const FrameParams* frame = nullptr);
// The camera rays of every view of a batch (RenderState::views), written to
// the views' pixels back to back.
void camera_rays_views(Ctx& ctx, uint32_t sample_number, uint32_t random_number, int32_t freeze, mipt_ray* rays,
                       hipStream_t stream, const FrameParams* frame);
void bounce_rays(Ctx& ctx, const mipt_ray* primary, const mipt_hit* hits, uint32_t* rng, mipt_ray* out,
                 uint32_t n, hipStream_t stream);

} // namespace mipt
