// api.cpp -- the extern "C" entry points of libmipt (declared in include/mipt.h).
// Each one validates its arguments, runs the C++ implementation and turns any
// failure into a negative status plus a message for mipt_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>

#include "internal.h"
#include "nrc.h"

namespace mipt {

void fail(int code, const std::string& msg) { throw Error{code, msg}; }

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        fail(e == hipErrorOutOfMemory ? MIPT_ERR_OUT_OF_MEMORY : MIPT_ERR_HIP,
             std::string(what) + ": " + hipGetErrorString(e));
}

void DeviceBuffer::release()
{
    if (ptr)
        (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

void DeviceBuffer::reserve(size_t n)
{
    if (n <= bytes && ptr)
        return;
    // Work already queued on any stream may still use the old allocation,
    // and hipFree does not wait for streams created non-blocking: wait for
    // the device before giving the memory back.
    if (ptr)
        HIP_CHECK(hipDeviceSynchronize());
    release();
    if (n == 0)
        n = 1;
    HIP_CHECK(hipMalloc(&ptr, n));
    bytes = n;
}

namespace {

template <class F> int guarded(Ctx* ctx, F&& f)
{
    try
    {
        if (ctx)
        {
            int cur = -1;
            if (hipGetDevice(&cur) == hipSuccess && cur != ctx->device)
                HIP_CHECK(hipSetDevice(ctx->device));
        }
        f();
        return MIPT_OK;
    }
    catch (const Error& e)
    {
        if (ctx)
            ctx->last_error = e.msg;
        return e.code;
    }
    catch (const std::bad_alloc&)
    {
        if (ctx)
            ctx->last_error = "host allocation failed";
        return MIPT_ERR_OUT_OF_MEMORY;
    }
}

template <class T> void upload(DeviceBuffer& buf, const T* host, size_t count)
{
    buf.reserve(sizeof(T) * (count > 0 ? count : 1));
    if (count > 0)
        HIP_CHECK(hipMemcpy(buf.ptr, host, sizeof(T) * count, hipMemcpyHostToDevice));
}

// Unpacks the CPU materials (CPUMaterial::pack_to_GPU, then the device
// record), uploads them and derives what the passes read from them: the
// lobe class of each material and whether some material transmits light or
// is dispersive.
void upload_materials(Ctx& ctx)
{
    RenderState& R = ctx.render;
    std::vector<float> mats(64 * (ctx.materials_host.size() > 0 ? ctx.materials_host.size() : 1), 0.0f);
    for (size_t i = 0; i < ctx.materials_host.size(); i++)
    {
        std::string why = unpack_material(ctx.materials_host[i], mats.data() + 64 * i);
        if (!why.empty())
            fail(MIPT_ERR_UNSUPPORTED, "material " + std::to_string(i) + " uses the " + why +
                                           ", which the path-tracing passes do not implement yet");
    }
    upload(R.materials, mats.data(), mats.size());
    R.scene_volumes = R.scene_dispersive = false;
    for (size_t i = 0; i < ctx.materials_host.size(); i++)
    {
        const float* o = mats.data() + 64 * i; // pt::Material: 22 specular / 23 diffuse transmission, 50 dispersion
        R.scene_volumes |= o[22] > 0.0f || o[23] > 0.0f;
        R.scene_dispersive |= o[50] > 0.0f && o[22] > 0.0f;
    }
    // Lobe class of each material: the set of Principled lobes with a
    // non-zero weight outside the object (Principled.h:1515-1554) and thin
    // film; materials of one class take the same branches in the BSDF.
    // Classes are numbered in order of first appearance; the 32nd and later
    // share class 31.
    {
        std::vector<int32_t> cls(ctx.materials_host.size() > 0 ? ctx.materials_host.size() : 1, 0);
        std::vector<uint32_t> seen;
        for (size_t i = 0; i < ctx.materials_host.size(); i++)
        {
            const float* o = mats.data() + 64 * i; // pt::Material floats
            const float metallic = o[7], srw = o[17], spec = o[19], st = o[22], dt = o[23];
            uint32_t mask = 0u;
            mask |= (o[20] > 0.0f) ? 1u : 0u;                                      // coat
            mask |= (o[21] > 0.0f) ? 2u : 0u;                                      // sheen
            mask |= (metallic > 0.0f && srw < 1.0f) ? 4u : 0u;                      // first metal
            mask |= (metallic > 0.0f && srw > 0.0f) ? 8u : 0u;                      // second metal
            mask |= ((1.0f - metallic) * (1.0f - st * (1.0f - dt)) * spec > 0.0f) ? 16u : 0u; // specular
            mask |= ((1.0f - metallic) * (1.0f - st) * (1.0f - dt) > 0.0f) ? 32u : 0u;    // diffuse
            mask |= ((1.0f - metallic) * (1.0f - dt) * st > 0.0f) ? 64u : 0u;             // glass
            mask |= ((1.0f - metallic) * dt > 0.0f) ? 128u : 0u;                            // diffuse transmission
            mask |= (o[24] > 0.0f) ? 256u : 0u;                                     // thin film
            size_t k = 0;
            while (k < seen.size() && seen[k] != mask)
                k++;
            if (k == seen.size() && seen.size() < 31)
                seen.push_back(mask);
            cls[i] = (int32_t)std::min<size_t>(k, 31);
        }
        R.lobe_classes = (int32_t)std::max<size_t>(seen.size(), 1);
        upload(R.mat_class, cls.data(), cls.size());
    }
}

// The renderer's table `t` (MIPT_LUT_GGX_CONDUCTOR .. MIPT_LUT_GGX_THIN_GLASS):
// its device buffer, element count and the shipped data.
struct LutSlot {
    DeviceBuffer* buf;
    size_t count;
    const float* shipped;
};
LutSlot lut_slot(RenderState& R, int32_t t)
{
    switch (t)
    {
    case MIPT_LUT_GGX_CONDUCTOR: return {&R.lut_conductor, 128 * 128, conductor_lut_data()};
    case MIPT_LUT_GLOSSY_DIELECTRIC: return {&R.lut_glossy, 128 * 64 * 128, glossy_dielectric_lut_data()};
    case MIPT_LUT_GGX_GLASS: return {&R.lut_glass, 256 * 16 * 128, glass_lut_data()};
    case MIPT_LUT_GGX_GLASS_INVERSE: return {&R.lut_glass_inv, 256 * 16 * 128, glass_inv_lut_data()};
    default: return {&R.lut_thin_glass, 32 * 32 * 96, thin_glass_lut_data()};
    }
}

void upload_lut(RenderState& R, int32_t t)
{
    const LutSlot sl = lut_slot(R, t);
    const std::vector<float>& o = R.lut_override[t];
    upload(*sl.buf, o.empty() ? sl.shipped : o.data(), sl.count);
}

// Scene data the passes need beyond the traversal: unpacked materials,
// triangle areas (ThreadFunctions.cpp:221-242: |cross(AB, AC)| * 0.5) and the
// energy-compensation LUT.
void prepare_scene_for_render(Ctx& ctx)
{
    RenderState& R = ctx.render;
    if (R.scene_ready)
        return;
    if (!ctx.scene_set)
        fail(MIPT_ERR_NO_SCENE, "no scene: call mipt_set_scene first");
    if (!ctx.bvh.valid)
        fail(MIPT_ERR_NO_BVH, "no BVH: call mipt_build_bvh first");
    if (ctx.triangle_count > 0 && (ctx.materials_host.empty() || !ctx.material_indices.ptr))
        fail(MIPT_ERR_INVALID_ARGUMENT, "the path-tracing passes need materials and material indices");
    upload_materials(ctx);
    std::vector<float> areas(ctx.host_vertices.size() ? ctx.triangle_count : 0);
    for (int32_t t = 0; t < (int32_t)areas.size(); t++)
    {
        const float* v[3];
        for (int k = 0; k < 3; k++)
            v[k] = ctx.host_vertices.data() + 3 * (size_t)ctx.host_indices[3 * (size_t)t + k];
        float ab[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
        float ac[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
        float n[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
        areas[t] = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) * 0.5f;
    }
    upload(R.areas, areas.data(), areas.size());
    for (int32_t t = MIPT_LUT_GGX_CONDUCTOR; t <= MIPT_LUT_GGX_THIN_GLASS; t++)
        upload_lut(R, t);
    upload(R.lut_glass_exponent, glass_exponent_data(), 2 * 10 * 11 * 4);
    upload(R.lut_sheen, sheen_ltc_data(), 32 * 32 * 3);
    build_power_alias_table(ctx);
    R.scene_ready = true;
}

void allocate_render_buffers(Ctx& ctx)
{
    RenderState& R = ctx.render;
    const size_t n = (size_t)R.settings.width * (size_t)R.settings.height;
    R.pixels = (uint32_t)n;
    R.frame_camera_valid = false;
    R.frame_count = 0;
    R.gb_prim.reserve(4 * n);
    R.gb_pos.reserve(16 * n);
    R.gb_gn.reserve(4 * n);
    R.gb_sn.reserve(4 * n);
    R.gb_mat.reserve(4 * n);
    for (DeviceBuffer* b : {&R.throughput, &R.thr_prev, &R.radiance, &R.vtx_pos, &R.vtx_sn, &R.vtx_gn, &R.vtx_view,
                            &R.nee_light, &R.nee_bsdf})
        b->reserve(16 * n);
    R.rng.reserve(4 * n);
    R.flags.reserve(4 * n);
    R.vol_a.reserve(16 * n);
    if (R.settings.direct_light_sampling_strategy == MIPT_LSS_RIS_BSDF_AND_LIGHT ||
        R.settings.direct_light_sampling_strategy == MIPT_LSS_RESTIR_DI)
        R.ris_state.reserve(48 * n);
    R.vol_b.reserve(16 * n);
    R.shadow_rays.reserve(32 * n);
    R.shadow_occ.reserve(n);
    R.bsdf_rays.reserve(32 * n);
    R.bsdf_hits.reserve(16 * n);
    R.cont_rays.reserve(32 * n);
    R.cont_hits.reserve(16 * n);
    const size_t qslots = (size_t)SEG_COUNT * seg_capacity((uint32_t)n);
    R.q_shadow.reserve(4 * qslots);
    R.q_bsdf.reserve(4 * qslots);
    R.q_cont.reserve(4 * qslots);
    R.q_skip.reserve(4 * 2 * qslots);
    R.q_live.reserve(4 * 2 * LIVE_APPENDERS * qslots);
    R.q_counts.reserve(4 * 7 * (size_t)SEG_COUNTER_WORDS);
    R.shade_sorted.reserve(4 * n);
    R.sort_counts.reserve(4 * 2 * (size_t)SHADE_SORT_COUNTERS * SEG_STRIDE);
    R.fb.reserve(12 * n);
    HIP_CHECK(hipMemset(R.fb.ptr, 0, 12 * n));
    if (R.settings.accumulate_fp64)
    {
        R.fb64.reserve(24 * n);
        HIP_CHECK(hipMemset(R.fb64.ptr, 0, 24 * n));
    }
    R.counters.reserve(64);
    HIP_CHECK(hipMemset(R.counters.ptr, 0, 64));
    // The G-buffer starts in its reset state: no hit, material -1.
    HIP_CHECK(hipMemset(R.gb_prim.ptr, 0xff, 4 * n));
    HIP_CHECK(hipMemset(R.gb_mat.ptr, 0xff, 4 * n));
    HIP_CHECK(hipMemset(R.gb_gn.ptr, 0, 4 * n));
    HIP_CHECK(hipMemset(R.gb_sn.ptr, 0, 4 * n));
    HIP_CHECK(hipMemset(R.gb_pos.ptr, 0, 16 * n));
    allocate_restir_buffers(ctx);
    // the shadow buffers hold one slot per pixel, more with envmap sampling
    R.shadow_slots = 1;
    allocate_env_path_buffers(ctx);
    if (restir_needs_prev_gbuffer(R))
    {
        HIP_CHECK(hipMemset(R.pg_prim.ptr, 0xff, 4 * n));
        HIP_CHECK(hipMemset(R.pg_mat.ptr, 0xff, 4 * n));
        HIP_CHECK(hipMemset(R.pg_gn.ptr, 0, 4 * n));
        HIP_CHECK(hipMemset(R.pg_sn.ptr, 0, 4 * n));
        HIP_CHECK(hipMemset(R.pg_pos.ptr, 0, 16 * n));
    }
    restir_reset(ctx, nullptr);
    R.last_output = 1;
    R.restir_shade_buffer = 1;
    R.temporal_output = 0;
    R.fused_output = 1;
    HIP_CHECK(hipDeviceSynchronize());
}

// Batched views render with the plain path tracer only: ReSTIR DI reuses
// neighbours in one image, and the radiance cache trains on one view's paths.
void check_views_settings(const mipt_render_settings& s)
{
    if (s.direct_light_sampling_strategy != MIPT_LSS_MIS_LIGHT_BSDF || s.radiance_cache)
        fail(MIPT_ERR_UNSUPPORTED, "batched views render with MIS light + BSDF sampling and no radiance cache");
}

uint64_t views_total(const std::vector<mipt_view>& views)
{
    uint64_t t = 0;
    for (const mipt_view& v : views)
        t += (uint64_t)v.width * (uint64_t)v.height;
    return t;
}

void destroy_graph(RenderState& R)
{
    if (R.graph_exec)
        (void)hipGraphExecDestroy(R.graph_exec);
    if (R.graph)
        (void)hipGraphDestroy(R.graph);
    R.graph_exec = nullptr;
    R.graph = nullptr;
}

// One plain sample (camera pass + path pass) as a replay of a captured
// hipGraph. The capture records the kernels with the device record of the
// per-sample values as their source; nothing is allocated while capturing
// (the traversal's spill stack is sized first; every other buffer is
// allocated with the render settings).
void render_sample_graph(Ctx& ctx, uint32_t sample_number, hipStream_t stream)
{
    RenderState& R = ctx.render;
    prepare_scene_for_render(ctx);
    const uint32_t k = sample_number * 2u;
    FrameParams p;
    p.sample_number = sample_number;
    p.rn_camera = mipt_host_random_number(k);
    p.rn_path = mipt_host_random_number(k + 1u);
    p.consistent_tau = consistent_tau(R.settings.microfacet_regularization_tau_0, sample_number);
    p.consistent_tau_refract = consistent_tau(R.settings.microfacet_regularization_tau_0, sample_number + 1);
    if (R.graph_dirty || !R.graph_exec)
    {
        destroy_graph(R);
        if (!R.capture_stream)
            HIP_CHECK(hipStreamCreateWithFlags(&R.capture_stream, hipStreamNonBlocking));
        R.frame_params.reserve(sizeof(FrameParams));
        trace_reserve(ctx);
        HIP_CHECK(hipDeviceSynchronize());
        const FrameParams* fp = R.frame_params.as<FrameParams>();
        HIP_CHECK(hipStreamBeginCapture(R.capture_stream, hipStreamCaptureModeRelaxed));
        try
        {
            camera_rays_pass(ctx, sample_number, p.rn_camera, R.capture_stream, fp);
            path_trace_pass(ctx, sample_number, p.rn_path, R.capture_stream, false, fp);
        }
        catch (...)
        {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(R.capture_stream, &g);
            if (g)
                (void)hipGraphDestroy(g);
            throw;
        }
        HIP_CHECK(hipStreamEndCapture(R.capture_stream, &R.graph));
        HIP_CHECK(hipGraphInstantiate(&R.graph_exec, R.graph, nullptr, nullptr, 0));
        R.graph_dirty = false;
        R.graph_replays = 0;
    }
    write_frame_params(R.frame_params.as<FrameParams>(), p, stream);
    HIP_CHECK(hipGraphLaunch(R.graph_exec, stream));
    R.graph_replays++;
}

} // namespace
} // namespace mipt

using mipt::Ctx;

namespace {
// Message of the last failure that had no context to hold it (mipt_create).
thread_local std::string g_no_ctx_error;

int create_fail(Ctx* ctx, const char* what, hipError_t e)
{
    g_no_ctx_error = std::string(what) + ": " + hipGetErrorString(e);
    delete ctx;
    return e == hipErrorOutOfMemory ? MIPT_ERR_OUT_OF_MEMORY : MIPT_ERR_HIP;
}
} // namespace

namespace mipt {
// Message of a failed call that has no context (mipt_last_error(NULL)).
void set_no_ctx_error(const std::string& msg) { g_no_ctx_error = msg; }
} // namespace mipt


extern "C" {

const char* mipt_version(void) { return "mipt 0.1 (gfx950)"; }

void mipt_material_default(mipt_material* m)
{
    if (!m)
        return;
    std::memset(m, 0, sizeof(*m));
    auto set3 = [](float* d, float v) { d[0] = d[1] = d[2] = v; };
    // Defaults of CPUMaterial, src/HostDeviceCommon/Material/MaterialCPU.h:213-330.
    set3(m->emission, 0.0f);
    m->emission_strength = 1.0f;
    m->global_emissive_factor = 1.0f;
    m->emissive_texture_used = 0;
    set3(m->base_color, 1.0f);
    m->roughness = 0.3f;
    m->oren_nayar_sigma = 0.34906585039886591538f;
    m->metallic = 0.0f;
    m->metallic_F90_falloff_exponent = 5.0f;
    set3(m->metallic_F82, 1.0f);
    set3(m->metallic_F90, 1.0f);
    m->anisotropy = 0.0f;
    m->anisotropy_rotation = 0.0f;
    m->second_roughness_weight = 0.0f;
    m->second_roughness = 0.5f;
    m->do_metallic_energy_compensation = 1;
    m->specular = 1.0f;
    m->specular_tint = 1.0f;
    set3(m->specular_color, 1.0f);
    m->specular_darkening = 0.0f;
    m->do_specular_energy_compensation = 1;
    m->coat = 0.0f;
    set3(m->coat_medium_absorption, 1.0f);
    m->coat_medium_thickness = 5.0f;
    m->coat_roughness = 0.0f;
    m->coat_roughening = 1.0f;
    m->coat_darkening = 1.0f;
    m->coat_anisotropy = 0.0f;
    m->coat_anisotropy_rotation = 0.0f;
    m->coat_ior = 1.5f;
    m->do_coat_energy_compensation = 1;
    m->sheen = 0.0f;
    m->sheen_roughness = 0.5f;
    set3(m->sheen_color, 1.0f);
    m->ior = 1.40f;
    m->specular_transmission = 0.0f;
    m->diffuse_transmission = 0.0f;
    m->absorption_at_distance = 1.0f;
    set3(m->absorption_color, 1.0f);
    m->dispersion_scale = 0.0f;
    m->dispersion_abbe_number = 20.0f;
    m->thin_walled = 0;
    m->do_glass_energy_compensation = 1;
    m->thin_film = 0.0f;
    m->thin_film_ior = 1.3f;
    m->thin_film_thickness = 500.0f;
    m->thin_film_kappa_3 = 0.0f;
    m->thin_film_hue_shift_degrees = 0.0f;
    m->thin_film_base_ior_override = 1.0f;
    m->thin_film_do_ior_override = 0;
    m->alpha_opacity = 1.0f;
    m->dielectric_priority = 0;
}

int mipt_create(int hip_device, mipt_ctx** out)
{
    if (!out)
        return MIPT_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    Ctx* ctx = nullptr;
    try
    {
        ctx = new Ctx;
    }
    catch (const std::bad_alloc&)
    {
        return MIPT_ERR_OUT_OF_MEMORY;
    }
    ctx->device = hip_device;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess)
        return create_fail(ctx, "hipGetDeviceCount", e);
    if (hip_device < 0 || hip_device >= count)
    {
        g_no_ctx_error = "device " + std::to_string(hip_device) + " out of range (" + std::to_string(count) +
                         " devices)";
        delete ctx;
        return MIPT_ERR_INVALID_ARGUMENT;
    }
    if ((e = hipSetDevice(hip_device)) != hipSuccess)
        return create_fail(ctx, "hipSetDevice", e);
    int cus = 0;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, hip_device)) != hipSuccess)
        return create_fail(ctx, "hipDeviceGetAttribute(MultiprocessorCount)", e);
    ctx->cu_count = cus;
    mipt::noalloc_scratch_init(hip_device);
    mipt_render_settings_default(&ctx->render.settings);
    mipt_restir_di_settings_default(&ctx->render.restir);
    mipt_world_settings_default(&ctx->world);
    *out = reinterpret_cast<mipt_ctx*>(ctx);
    return MIPT_OK;
}

void mipt_destroy(mipt_ctx* c)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx)
        return;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    mipt::destroy_graph(ctx->render);
    if (ctx->render.capture_stream)
        (void)hipStreamDestroy(ctx->render.capture_stream);
    delete ctx;
}

const char* mipt_last_error(const mipt_ctx* c)
{
    const Ctx* ctx = reinterpret_cast<const Ctx*>(c);
    return ctx ? ctx->last_error.c_str() : g_no_ctx_error.c_str();
}

int mipt_set_scene(mipt_ctx* c, const mipt_scene_desc* s)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (ctx)
        ctx->render.graph_dirty = true; // a captured sample would replay stale state
    if (!ctx)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        if (!s || s->triangle_count < 0 || s->vertex_count < 0 || s->material_count < 0 ||
            s->emissive_triangle_count < 0)
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "mipt_set_scene: negative counts or null scene");
        if (s->triangle_count > 0 && (!s->triangles_vertex_indices || !s->vertices_positions))
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "mipt_set_scene: missing geometry arrays");
        if (s->triangle_count >= (1 << 27))
            mipt::fail(MIPT_ERR_UNSUPPORTED, "mipt_set_scene: more than 2^27 triangles");
        for (int64_t i = 0; i < 3 * (int64_t)s->triangle_count; i++)
        {
            int32_t v = s->triangles_vertex_indices[i];
            if (v < 0 || v >= s->vertex_count)
                mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "mipt_set_scene: vertex index out of range at " +
                                                          std::to_string(i));
        }
        if (s->material_indices)
            for (int32_t i = 0; i < s->triangle_count; i++)
                if (s->material_indices[i] < 0 || s->material_indices[i] >= s->material_count)
                    mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "mipt_set_scene: material index out of range");
        for (int32_t i = 0; i < s->emissive_triangle_count; i++)
            if (s->emissive_triangles_primitive_indices[i] < 0 ||
                s->emissive_triangles_primitive_indices[i] >= s->triangle_count)
                mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "mipt_set_scene: emissive index out of range");
        ctx->scene_set = false;
        ctx->bvh.valid = false;
        mipt::upload(ctx->vertices, s->vertices_positions, 3 * (size_t)s->vertex_count);
        mipt::upload(ctx->indices, s->triangles_vertex_indices, 3 * (size_t)s->triangle_count);
        if (s->has_vertex_normals && s->vertex_normals)
        {
            mipt::upload(ctx->has_normals, s->has_vertex_normals, (size_t)s->vertex_count);
            mipt::upload(ctx->vertex_normals, s->vertex_normals, 3 * (size_t)s->vertex_count);
        }
        else
            ctx->has_normals.release(), ctx->vertex_normals.release();
        if (s->material_indices)
            mipt::upload(ctx->material_indices, s->material_indices, (size_t)s->triangle_count);
        else
            ctx->material_indices.release();
        mipt::upload(ctx->emissive_prims, s->emissive_triangles_primitive_indices, (size_t)s->emissive_triangle_count);
        ctx->materials_host.assign(s->materials, s->materials + (s->materials ? s->material_count : 0));
        ctx->host_vertices.assign(s->vertices_positions, s->vertices_positions + 3 * (size_t)s->vertex_count);
        ctx->host_indices.assign(s->triangles_vertex_indices,
                                 s->triangles_vertex_indices + 3 * (size_t)s->triangle_count);
        ctx->render.scene_ready = false;
        for (int a = 0; a < 3; a++)
        {
            ctx->scene_bmin[a] = s->vertex_count > 0 ? s->vertices_positions[a] : 0.0f;
            ctx->scene_bmax[a] = ctx->scene_bmin[a];
        }
        for (int32_t v = 0; v < s->vertex_count; v++)
            for (int a = 0; a < 3; a++)
            {
                const float x = s->vertices_positions[3 * (size_t)v + a];
                ctx->scene_bmin[a] = x < ctx->scene_bmin[a] ? x : ctx->scene_bmin[a];
                ctx->scene_bmax[a] = x > ctx->scene_bmax[a] ? x : ctx->scene_bmax[a];
            }
        ctx->triangle_count = s->triangle_count;
        ctx->vertex_count = s->vertex_count;
        ctx->material_count = s->material_count;
        ctx->emissive_count = s->emissive_triangle_count;
        ctx->scene_set = true;
    });
}

int mipt_build_bvh(mipt_ctx* c, int32_t builder)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (ctx)
        ctx->render.graph_dirty = true; // a captured sample would replay stale state
    if (!ctx)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        if (!ctx->scene_set)
            mipt::fail(MIPT_ERR_NO_SCENE, "mipt_build_bvh: no scene");
        if (builder == MIPT_BVH_LBVH)
            mipt::build_lbvh(*ctx);
        else if (builder == MIPT_BVH_SAH_BINNED)
            mipt::build_sah_binned(*ctx);
        else
            mipt::fail(MIPT_ERR_UNSUPPORTED, "mipt_build_bvh: unknown builder " + std::to_string(builder));
    });
}

int mipt_get_bvh_info(const mipt_ctx* c, mipt_bvh_info* out)
{
    const Ctx* ctx = reinterpret_cast<const Ctx*>(c);
    if (!ctx || !out)
        return MIPT_ERR_INVALID_ARGUMENT;
    if (!ctx->bvh.valid)
        return MIPT_ERR_NO_BVH;
    const mipt::Bvh& b = ctx->bvh;
    out->builder = b.builder;
    out->triangle_count = b.triangle_count;
    out->node_count = b.node_count;
    out->leaf_count = b.leaf_count;
    out->max_depth = b.max_depth;
    out->max_leaf_size = b.max_leaf_size;
    out->node_bytes = (int64_t)b.node_count * (int64_t)sizeof(mipt::BvhNode);
    out->triangle_bytes = (int64_t)b.triangle_count * (int64_t)sizeof(mipt::TriRecord);
    out->build_ms = b.build_ms;
    out->sah_cost = b.sah_cost;
    return MIPT_OK;
}

int mipt_read_bvh(const mipt_ctx* c, void* host_nodes, void* host_triangles)
{
    const Ctx* ctx = reinterpret_cast<const Ctx*>(c);
    if (!ctx)
        return MIPT_ERR_INVALID_ARGUMENT;
    if (!ctx->bvh.valid)
        return MIPT_ERR_NO_BVH;
    const mipt::Bvh& b = ctx->bvh;
    if (host_nodes && b.node_count > 0)
        if (hipMemcpy(host_nodes, b.nodes.ptr, sizeof(mipt::BvhNode) * (size_t)b.node_count,
                      hipMemcpyDeviceToHost) != hipSuccess)
            return MIPT_ERR_HIP;
    if (host_triangles && b.triangle_count > 0)
        if (hipMemcpy(host_triangles, b.tris.ptr, sizeof(mipt::TriRecord) * (size_t)b.triangle_count,
                      hipMemcpyDeviceToHost) != hipSuccess)
            return MIPT_ERR_HIP;
    return MIPT_OK;
}

int mipt_set_trace_config(mipt_ctx* c, int32_t block_size, int32_t lds_stack)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (ctx)
        ctx->render.graph_dirty = true; // a captured sample would replay stale state
    if (!ctx)
        return MIPT_ERR_INVALID_ARGUMENT;
    if (block_size != 0 && block_size != 64 && block_size != 128 && block_size != 256)
        return MIPT_ERR_INVALID_ARGUMENT;
    if (lds_stack != 0 && lds_stack != 8 && lds_stack != 16 && lds_stack != 24 && lds_stack != 32)
        return MIPT_ERR_INVALID_ARGUMENT;
    if (block_size)
        ctx->trace_cfg.block_size = block_size;
    if (lds_stack)
        ctx->trace_cfg.lds_stack = lds_stack;
    return MIPT_OK;
}

int mipt_update_materials(mipt_ctx* c, const mipt_material* materials, int32_t count)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (ctx)
        ctx->render.graph_dirty = true; // a captured sample would replay stale state
    if (!ctx || !materials)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        if (!ctx->scene_set)
            mipt::fail(MIPT_ERR_NO_SCENE, "mipt_update_materials: no scene");
        if (count != (int32_t)ctx->materials_host.size())
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "mipt_update_materials: " + std::to_string(count) +
                                                      " materials for a scene of " +
                                                      std::to_string(ctx->materials_host.size()));
        const std::vector<mipt_material> old = ctx->materials_host;
        ctx->materials_host.assign(materials, materials + count);
        if (ctx->render.scene_ready)
        {
            // kernels still queued on any stream may read the material records
            HIP_CHECK(hipDeviceSynchronize());
            try
            {
                mipt::upload_materials(*ctx);
            }
            catch (...)
            {
                ctx->materials_host = old; // an unsupported material leaves the scene as it was
                mipt::upload_materials(*ctx);
                throw;
            }
        }
    });
}

int mipt_update_material(mipt_ctx* c, int32_t index, const mipt_material* material)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || !material)
        return MIPT_ERR_INVALID_ARGUMENT;
    if (index < 0 || index >= (int32_t)ctx->materials_host.size())
    {
        ctx->last_error = "mipt_update_material: index " + std::to_string(index) + " out of range";
        return MIPT_ERR_INVALID_ARGUMENT;
    }
    std::vector<mipt_material> all = ctx->materials_host;
    all[(size_t)index] = *material;
    return mipt_update_materials(c, all.data(), (int32_t)all.size());
}

int mipt_set_shading_order(mipt_ctx* c, int32_t order)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (ctx)
        ctx->render.graph_dirty = true; // a captured sample would replay stale state
    if (!ctx || order < MIPT_SHADE_ORDER_LIVE || order > MIPT_SHADE_ORDER_AUTO)
        return MIPT_ERR_INVALID_ARGUMENT;
    ctx->render.shade_order = order;
    return MIPT_OK;
}

int mipt_trace_closest(mipt_ctx* c, const mipt_ray* d_rays, mipt_hit* d_hits, uint32_t n, void* stream)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || (n > 0 && (!d_rays || !d_hits)))
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        mipt::trace_closest(*ctx, d_rays, d_hits, nullptr, n, reinterpret_cast<hipStream_t>(stream));
    });
}

int mipt_trace_closest_noalloc(mipt_ctx* c, const mipt_ray* d_rays, mipt_hit* d_hits, uint32_t n, void* stream)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || (n > 0 && (!d_rays || !d_hits)))
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        mipt::trace_closest_noalloc(*ctx, d_rays, d_hits, n, reinterpret_cast<hipStream_t>(stream));
    });
}

int mipt_trace_closest_counted(mipt_ctx* c, const mipt_ray* d_rays, mipt_hit* d_hits, uint32_t* d_counts, uint32_t n,
                               void* stream)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || (n > 0 && (!d_rays || !d_hits || !d_counts)))
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        mipt::trace_closest(*ctx, d_rays, d_hits, d_counts, n, reinterpret_cast<hipStream_t>(stream));
    });
}

int mipt_trace_any(mipt_ctx* c, const mipt_ray* d_rays, uint8_t* d_occluded, uint32_t n, void* stream)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || (n > 0 && (!d_rays || !d_occluded)))
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        mipt::trace_any(*ctx, d_rays, d_occluded, n, reinterpret_cast<hipStream_t>(stream));
    });
}

int mipt_camera_rays(mipt_ctx* c, const mipt_camera* cam, int32_t w, int32_t h, uint32_t sample_number,
                     uint32_t random_number, int32_t freeze_random, mipt_ray* d_rays, uint32_t* d_rng_state,
                     void* stream)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || !cam || w <= 0 || h <= 0 || !d_rays)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        mipt::camera_rays(*ctx, *cam, w, h, sample_number, random_number, freeze_random, d_rays, d_rng_state,
                          reinterpret_cast<hipStream_t>(stream));
    });
}

int mipt_bounce_rays(mipt_ctx* c, const mipt_ray* d_primary, const mipt_hit* d_hits, uint32_t* d_rng_state,
                     mipt_ray* d_out, uint32_t n, void* stream)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || (n > 0 && (!d_primary || !d_hits || !d_rng_state || !d_out)))
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        mipt::bounce_rays(*ctx, d_primary, d_hits, d_rng_state, d_out, n, reinterpret_cast<hipStream_t>(stream));
    });
}

void mipt_render_settings_default(mipt_render_settings* s)
{
    if (!s)
        return;
    std::memset(s, 0, sizeof(*s));
    s->width = 512;
    s->height = 512;
    s->nb_bounces = 4;
    s->do_russian_roulette = 1;
    // RenderSettings.h:107 evaluates min(5, nb_bounces / 2) while nb_bounces
    // is still 0, so Russian roulette starts at bounce 0.
    s->russian_roulette_min_depth = 0;
    s->russian_roulette_throughput_clamp = 10.0f;
    s->freeze_random = 0;
    s->enable_direct = 1;
    s->microfacet_regularization_tau_0 = 30.0f;
    s->microfacet_regularization_min_roughness = 0.0f;
    s->do_microfacet_regularization = 1;
    s->accumulate_fp64 = 1;
    s->direct_light_sampling_strategy = MIPT_LSS_MIS_LIGHT_BSDF;
    s->direct_light_sampling_base_strategy = MIPT_LSS_BASE_UNIFORM;
    s->do_normal_mapping = 1;
    s->accumulate = 1;
    s->ris_number_of_light_candidates = 4; // RISSettings.h:12-15
    s->ris_number_of_bsdf_candidates = 1;
}

void mipt_restir_di_settings_default(mipt_restir_di_settings* s)
{
    if (!s)
        return;
    std::memset(s, 0, sizeof(*s));
    // ReSTIRDISettings() (ReSTIRDISettings.h:58-121) and ReSTIRDIInitialCandidatesSettings (:17-19)
    s->number_of_initial_light_candidates = 4;
    s->number_of_initial_bsdf_candidates = 1;
    s->do_temporal_reuse_pass = 1;
    s->do_spatial_reuse_pass = 1;
    s->spatial_number_of_passes = 1;
    s->spatial_reuse_radius = 16;
    s->spatial_reuse_neighbor_count = 5;
    s->m_cap = 3;
    s->use_confidence_weights = 1;
    s->use_normal_similarity_heuristic = 1;
    s->normal_similarity_angle_precomp = 0.906307787f;
    s->use_plane_distance_heuristic = 1;
    s->plane_distance_threshold = 0.1f;
    s->reject_using_geometric_normals = 1;
    s->do_final_shading_visibility = 1;
    s->do_fused_spatiotemporal = 0;
    s->do_light_presampling = 0;
    // ReSTIRDILightPresamplingSettings (ReSTIRDISettings.h:45-56)
    s->light_presampling_number_of_subsets = 128;
    s->light_presampling_subset_size = 1024;
    s->light_presampling_tile_size = 8;
    s->envmap_candidate_probability = 0.5f; // ReSTIRDISettings.h:22
    // ReSTIRCommonTemporalPassSettings (ReSTIRDISettings.h:65-69)
    s->use_permutation_sampling = 0;
    s->permutation_sampling_random_bits = 42;
    s->max_neighbor_search_count = 8;
    s->neighbor_search_radius = 4;
    s->later_bounces_sampling_strategy = MIPT_RESTIR_DI_LATER_BOUNCES_RIS_BSDF_AND_LIGHT; // ReSTIRDIOptions.h:140
}

int mipt_set_restir_di_settings(mipt_ctx* c, const mipt_restir_di_settings* s)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (ctx)
        ctx->render.graph_dirty = true; // a captured sample would replay stale state
    if (!ctx || !s || s->number_of_initial_light_candidates < 0 || s->number_of_initial_bsdf_candidates < 0 ||
        s->spatial_number_of_passes < 1 || s->spatial_reuse_radius < 0 || s->spatial_reuse_neighbor_count < 0 ||
        s->spatial_reuse_neighbor_count > 32 || s->m_cap < 0 ||
        !(s->envmap_candidate_probability >= 0.0f && s->envmap_candidate_probability <= 1.0f) ||
        (s->do_light_presampling &&
         (s->light_presampling_number_of_subsets < 1 || s->light_presampling_subset_size < 1 ||
          s->light_presampling_tile_size < 1 ||
          (int64_t)s->light_presampling_number_of_subsets * s->light_presampling_subset_size > (1ll << 24))) ||
        s->max_neighbor_search_count < 0 || s->max_neighbor_search_count > 64 || s->neighbor_search_radius < 0 ||
        s->neighbor_search_radius > 4096 ||
        (s->later_bounces_sampling_strategy != MIPT_RESTIR_DI_LATER_BOUNCES_MIS_LIGHT_BSDF &&
         s->later_bounces_sampling_strategy != MIPT_RESTIR_DI_LATER_BOUNCES_RIS_BSDF_AND_LIGHT))
        return MIPT_ERR_INVALID_ARGUMENT;
    if (s->do_fused_spatiotemporal && !s->do_temporal_reuse_pass)
    {
        ctx->last_error = "the fused spatiotemporal pass needs do_temporal_reuse_pass (it reads last frame's "
                          "reservoirs and G-buffer)";
        return MIPT_ERR_UNSUPPORTED;
    }
    return mipt::guarded(ctx, [&] {
        // New permutation bits alone (the per-frame value of the temporal
        // passes) keep the buffers and the reservoirs in them.
        mipt_restir_di_settings same = *s;
        same.permutation_sampling_random_bits = ctx->render.restir.permutation_sampling_random_bits;
        const bool bits_only = std::memcmp(&same, &ctx->render.restir, sizeof(same)) == 0;
        ctx->render.restir = *s;
        if (ctx->render.configured && !bits_only)
            mipt::allocate_render_buffers(*ctx);
    });
}

int mipt_set_render_settings(mipt_ctx* c, const mipt_render_settings* s)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (ctx)
        ctx->render.graph_dirty = true; // a captured sample would replay stale state
    if (!ctx || !s || s->width <= 0 || s->height <= 0 || s->nb_bounces < 0 ||
        (int64_t)s->width * (int64_t)s->height > (1ll << 28) ||
        (s->direct_light_sampling_strategy != MIPT_LSS_MIS_LIGHT_BSDF &&
         s->direct_light_sampling_strategy != MIPT_LSS_RIS_BSDF_AND_LIGHT &&
         s->direct_light_sampling_strategy != MIPT_LSS_RESTIR_DI) ||
        s->ris_number_of_light_candidates < 0 || s->ris_number_of_light_candidates > 64 ||
        s->ris_number_of_bsdf_candidates < 0 || s->ris_number_of_bsdf_candidates > 1 ||
        (s->direct_light_sampling_base_strategy != MIPT_LSS_BASE_UNIFORM &&
         s->direct_light_sampling_base_strategy != MIPT_LSS_BASE_POWER) ||
        (s->radiance_cache != 0 && s->radiance_cache != 1))
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        mipt_render_settings settings = *s;
        if (!ctx->render.views.empty())
        {
            mipt::check_views_settings(settings);
            settings.width = (int32_t)mipt::views_total(ctx->render.views);
            settings.height = 1;
        }
        ctx->render.settings = settings;
        ctx->render.configured = false;
        mipt::prepare_scene_for_render(*ctx);
        mipt::allocate_render_buffers(*ctx);
        ctx->render.configured = true;
    });
}

int mipt_set_views(mipt_ctx* c, const mipt_view* views, int32_t count)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || count < 0 || (count > 0 && !views))
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        mipt::RenderState& R = ctx->render;
        R.graph_dirty = true;
        if (count == 0)
        {
            R.views.clear();
            R.configured = false; // mipt_set_render_settings sets the single view's size again
            return;
        }
        uint64_t total = 0;
        for (int32_t v = 0; v < count; v++)
        {
            if (views[v].width <= 0 || views[v].height <= 0)
                mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "mipt_set_views: view " + std::to_string(v) + " has no pixels");
            total += (uint64_t)views[v].width * (uint64_t)views[v].height;
        }
        if (total > (1ull << 28))
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "mipt_set_views: more than 2^28 pixels in all");
        mipt_render_settings settings = R.settings;
        mipt::check_views_settings(settings);
        std::vector<mipt::ViewDev> table((size_t)count);
        uint32_t start = 0;
        for (int32_t v = 0; v < count; v++)
        {
            mipt::ViewDev& d = table[(size_t)v];
            std::memcpy(d.iv, views[v].camera.inverse_view, sizeof(d.iv));
            std::memcpy(d.ip, views[v].camera.inverse_projection, sizeof(d.ip));
            std::memcpy(d.position, views[v].camera.position, sizeof(d.position));
            d.jitter = views[v].camera.do_jittering;
            d.width = views[v].width;
            d.height = views[v].height;
            d.start = start;
            start += (uint32_t)views[v].width * (uint32_t)views[v].height;
        }
        HIP_CHECK(hipDeviceSynchronize()); // queued kernels may read the old table
        mipt::upload(R.view_table, table.data(), table.size());
        R.views.assign(views, views + count);
        settings.width = (int32_t)total;
        settings.height = 1;
        R.settings = settings;
        R.configured = false;
        mipt::prepare_scene_for_render(*ctx);
        mipt::allocate_render_buffers(*ctx);
        R.configured = true;
        R.camera = views[0].camera; // the first view's camera (ReSTIR DI and the cache are off)
        R.camera_set = true;
    });
}

int mipt_set_graph_capture(mipt_ctx* c, int32_t enable)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || (enable != 0 && enable != 1))
        return MIPT_ERR_INVALID_ARGUMENT;
    ctx->render.graph_enabled = enable != 0;
    ctx->render.graph_dirty = true;
    return MIPT_OK;
}

int mipt_graph_replays(const mipt_ctx* c, uint32_t* out)
{
    const Ctx* ctx = reinterpret_cast<const Ctx*>(c);
    if (!ctx || !out)
        return MIPT_ERR_INVALID_ARGUMENT;
    *out = ctx->render.graph_replays;
    return MIPT_OK;
}

int mipt_set_queue_tile(mipt_ctx* c, int32_t tile)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || (tile != 1 && tile != 2 && tile != 4))
        return MIPT_ERR_INVALID_ARGUMENT;
    ctx->trace_cfg.queue_tile = tile;
    ctx->render.graph_dirty = true;
    return MIPT_OK;
}

int mipt_autotune_trace(mipt_ctx* c, const mipt_ray* d_rays, uint32_t n, int32_t repeats, void* stream,
                        mipt_trace_tuning* out, double* table_ms)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || !d_rays || n == 0 || repeats < 1)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        ctx->render.graph_dirty = true;
        mipt::autotune_trace(*ctx, d_rays, n, repeats, reinterpret_cast<hipStream_t>(stream), out, table_ms);
    });
}

int mipt_set_camera(mipt_ctx* c, const mipt_camera* cam)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (ctx)
        ctx->render.graph_dirty = true; // a captured sample would replay stale state
    if (!ctx || !cam)
        return MIPT_ERR_INVALID_ARGUMENT;
    ctx->render.camera = *cam;
    ctx->render.camera_set = true;
    return MIPT_OK;
}

int mipt_launch_pass(mipt_ctx* c, int32_t pass, uint32_t sample_number, uint32_t random_number, void* stream)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    const bool spatial = pass >= MIPT_PASS_RESTIR_DI_SPATIAL_REUSE(0) && pass < MIPT_PASS_RESTIR_DI_SPATIAL_REUSE(64);
    if (!ctx || (pass != MIPT_PASS_CAMERA_RAYS && pass != MIPT_PASS_PATH_TRACE &&
                 pass != MIPT_PASS_RESTIR_DI_INITIAL_CANDIDATES && pass != MIPT_PASS_RESTIR_DI_TEMPORAL_REUSE &&
                 pass != MIPT_PASS_RADIANCE_CACHE_TRAIN && pass != MIPT_PASS_RADIANCE_CACHE_WARMUP &&
                 pass != MIPT_PASS_RESTIR_DI_LIGHTS_PRESAMPLING && pass != MIPT_PASS_RESTIR_DI_SPATIOTEMPORAL_REUSE &&
                 !spatial))
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        if (!ctx->render.configured)
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "mipt_launch_pass: call mipt_set_render_settings first");
        if (!ctx->render.camera_set)
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "mipt_launch_pass: call mipt_set_camera first");
        mipt::prepare_scene_for_render(*ctx);
        // Not accumulating, every frame is sample 0 (MegaKernelRenderPass.cpp:42-43).
        if (!ctx->render.settings.accumulate)
            sample_number = 0;
        // The ReSTIR DI passes rebuild each surface's volume state, which for
        // a dispersive glass draws a wavelength from the pass's stream at
        // every surface fetch (ReSTIR/Surface.h:25-43); those draws are not
        // implemented.
        if (ctx->render.scene_dispersive && ctx->render.settings.direct_light_sampling_strategy == MIPT_LSS_RESTIR_DI)
            mipt::fail(MIPT_ERR_UNSUPPORTED, "ReSTIR DI with a dispersive glass material (dispersion_scale > 0 and "
                                             "specular_transmission > 0) is not implemented yet");
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        if (pass == MIPT_PASS_CAMERA_RAYS)
            mipt::camera_rays_pass(*ctx, sample_number, random_number, s);
        else if (pass == MIPT_PASS_PATH_TRACE)
            mipt::path_trace_pass(*ctx, sample_number, random_number, s);
        else if (pass == MIPT_PASS_RESTIR_DI_INITIAL_CANDIDATES)
            mipt::restir_initial_pass(*ctx, sample_number, random_number, s);
        else if (pass == MIPT_PASS_RESTIR_DI_TEMPORAL_REUSE)
            mipt::restir_temporal_pass(*ctx, sample_number, random_number, s);
        else if (pass == MIPT_PASS_RESTIR_DI_LIGHTS_PRESAMPLING)
            mipt::restir_presampling_pass(*ctx, sample_number, random_number, s);
        else if (pass == MIPT_PASS_RESTIR_DI_SPATIOTEMPORAL_REUSE)
            mipt::restir_spatiotemporal_pass(*ctx, sample_number, random_number, s);
        else if (pass == MIPT_PASS_RADIANCE_CACHE_TRAIN)
            mipt::nrc_train_pass(*ctx, s);
        else if (pass == MIPT_PASS_RADIANCE_CACHE_WARMUP)
            mipt::path_trace_pass(*ctx, sample_number, random_number, s, true);
        else
            mipt::restir_spatial_pass(*ctx, pass - MIPT_PASS_RESTIR_DI_SPATIAL_REUSE(0), sample_number, random_number,
                                      s);
    });
}

int32_t mipt_draws_per_sample(const mipt_ctx* c)
{
    const Ctx* ctx = reinterpret_cast<const Ctx*>(c);
    if (!ctx)
        return MIPT_ERR_INVALID_ARGUMENT;
    const mipt::RenderState& R = ctx->render;
    int32_t draws = 2;
    if (R.settings.direct_light_sampling_strategy == MIPT_LSS_RESTIR_DI)
    {
        draws += 1;
        if (R.restir.do_fused_spatiotemporal)
            draws += 3 + (R.restir.spatial_number_of_passes - 1);
        else
        {
            if (R.restir.do_temporal_reuse_pass)
                draws += 2;
            if (R.restir.do_spatial_reuse_pass)
                draws += R.restir.spatial_number_of_passes;
        }
    }
    return draws;
}

namespace {
// The passes of one sample; a warm-up frame runs RADIANCE_CACHE_WARMUP in
// place of PATH_TRACE.
int render_sample_passes(mipt_ctx* c, uint32_t sample_number, void* stream, bool warmup)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx)
        return MIPT_ERR_INVALID_ARGUMENT;
    if (warmup && !ctx->render.settings.radiance_cache)
        return MIPT_ERR_INVALID_ARGUMENT;
    {
        mipt::RenderState& R = ctx->render;
        const bool plain = R.settings.direct_light_sampling_strategy == MIPT_LSS_MIS_LIGHT_BSDF &&
                           !R.settings.radiance_cache && !warmup;
        if (R.graph_enabled && plain && sample_number > 0 && R.configured && R.camera_set && R.settings.accumulate)
            return mipt::guarded(ctx, [&] {
                mipt::render_sample_graph(*ctx, sample_number, reinterpret_cast<hipStream_t>(stream));
            });
        R.graph_replays = 0;
    }
    const uint32_t draws = (uint32_t)mipt_draws_per_sample(c);
    // Not accumulating, the host stream runs on from frame to frame
    // (GPURenderer::reset leaves it alone then).
    uint32_t k = (ctx->render.settings.accumulate ? sample_number : ctx->render.frame_count++) * draws;
    int rc = mipt_launch_pass(c, MIPT_PASS_CAMERA_RAYS, sample_number, mipt_host_random_number(k++), stream);
    const mipt::RenderState& R = ctx->render;
    if (rc == MIPT_OK && R.settings.direct_light_sampling_strategy == MIPT_LSS_RESTIR_DI)
    {
        // ReSTIRDIRenderPass::launch_async (ReSTIRDIRenderPass.cpp:343-381):
        // the presampling kernel reads the camera pass's random number
        if (R.restir.do_light_presampling)
            rc = mipt_launch_pass(c, MIPT_PASS_RESTIR_DI_LIGHTS_PRESAMPLING, sample_number,
                                  mipt_host_random_number(k - 1), stream);
        if (rc == MIPT_OK)
            rc = mipt_launch_pass(c, MIPT_PASS_RESTIR_DI_INITIAL_CANDIDATES, sample_number,
                                  mipt_host_random_number(k++), stream);
        if (rc == MIPT_OK && R.restir.do_fused_spatiotemporal)
        {
            // the temporal pass's number and permutation bits are drawn, then
            // the fused kernel's (configure_temporal_pass_for_fused_spatiotemporal,
            // configure_spatial_pass)
            ctx->render.restir.permutation_sampling_random_bits = (int32_t)mipt_host_random_number(k + 1);
            k += 2;
            rc = mipt_launch_pass(c, MIPT_PASS_RESTIR_DI_SPATIOTEMPORAL_REUSE, sample_number,
                                  mipt_host_random_number(k++), stream);
            for (int32_t i = 1; rc == MIPT_OK && i < R.restir.spatial_number_of_passes; i++)
                rc = mipt_launch_pass(c, MIPT_PASS_RESTIR_DI_SPATIAL_REUSE(i), sample_number,
                                      mipt_host_random_number(k++), stream);
        }
        else
        {
            if (rc == MIPT_OK && R.restir.do_temporal_reuse_pass)
            {
                // the second draw is the permutation-sampling bits (configure_temporal_pass)
                ctx->render.restir.permutation_sampling_random_bits = (int32_t)mipt_host_random_number(k + 1);
                rc = mipt_launch_pass(c, MIPT_PASS_RESTIR_DI_TEMPORAL_REUSE, sample_number,
                                      mipt_host_random_number(k), stream);
                k += 2;
            }
            for (int32_t i = 0;
                 rc == MIPT_OK && R.restir.do_spatial_reuse_pass && i < R.restir.spatial_number_of_passes; i++)
                rc = mipt_launch_pass(c, MIPT_PASS_RESTIR_DI_SPATIAL_REUSE(i), sample_number,
                                      mipt_host_random_number(k++), stream);
        }
    }
    if (rc != MIPT_OK)
        return rc;
    rc = mipt_launch_pass(c, warmup ? MIPT_PASS_RADIANCE_CACHE_WARMUP : MIPT_PASS_PATH_TRACE, sample_number,
                          mipt_host_random_number(k), stream);
    if (rc == MIPT_OK && R.settings.radiance_cache)
        rc = mipt_launch_pass(c, MIPT_PASS_RADIANCE_CACHE_TRAIN, sample_number, 0u, stream);
    return rc;
}
} // namespace

int mipt_render_sample(mipt_ctx* c, uint32_t sample_number, void* stream)
{
    return render_sample_passes(c, sample_number, stream, false);
}

int mipt_radiance_cache_warmup(mipt_ctx* c, uint32_t sample_number, void* stream)
{
    return render_sample_passes(c, sample_number, stream, true);
}

int mipt_get_power_alias_table(mipt_ctx* c, float* probas, int32_t* alias, float* sum)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        if (!ctx->render.scene_ready)
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "no alias table: call mipt_set_render_settings first");
        const mipt::RenderState& R = ctx->render;
        if (probas)
            std::memcpy(probas, R.alias_probas_host.data(), sizeof(float) * R.alias_probas_host.size());
        if (alias)
            std::memcpy(alias, R.alias_alias_host.data(), sizeof(int32_t) * R.alias_alias_host.size());
        if (sum)
            *sum = R.alias_sum;
    });
}

void mipt_world_settings_default(mipt_world_settings* w)
{
    if (!w)
        return;
    std::memset(w, 0, sizeof(*w));
    w->ambient_light_type = MIPT_AMBIENT_UNIFORM;
    w->envmap_intensity = 1.0f;
    for (int k = 0; k < 3; k++)
    {
        w->envmap_to_world[k][k] = 1.0f;
        w->world_to_envmap[k][k] = 1.0f;
    }
}

int mipt_set_world_settings(mipt_ctx* c, const mipt_world_settings* w)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || !w || w->ambient_light_type < MIPT_AMBIENT_NONE || w->ambient_light_type > MIPT_AMBIENT_ENVMAP ||
        !(w->envmap_intensity >= 0.0f))
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        if (w->ambient_light_type == MIPT_AMBIENT_ENVMAP && ctx->env_width == 0)
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "MIPT_AMBIENT_ENVMAP: call mipt_set_envmap first");
        HIP_CHECK(hipDeviceSynchronize()); // queued passes read the previous settings by value only
        ctx->world = *w;
        ctx->render.graph_dirty = true;
        mipt::allocate_env_path_buffers(*ctx);
    });
}

int mipt_get_world_settings(const mipt_ctx* c, mipt_world_settings* out)
{
    const Ctx* ctx = reinterpret_cast<const Ctx*>(c);
    if (!ctx || !out)
        return MIPT_ERR_INVALID_ARGUMENT;
    *out = ctx->world;
    return MIPT_OK;
}

int mipt_set_envmap(mipt_ctx* c, const float* pixels, int32_t width, int32_t height, int32_t channels)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        HIP_CHECK(hipDeviceSynchronize()); // queued passes may read the old texels
        ctx->render.graph_dirty = true;
        mipt::set_envmap(*ctx, pixels, width, height, channels);
    });
}

int mipt_get_envmap_tables(const mipt_ctx* c, float* probas, int32_t* alias, float* luminance_sum,
                           uint32_t* packed_texels)
{
    const Ctx* ctx = reinterpret_cast<const Ctx*>(c);
    if (!ctx)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(const_cast<Ctx*>(ctx), [&] {
        if (ctx->env_width == 0)
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "no envmap: call mipt_set_envmap first");
        const size_t count = (size_t)ctx->env_width * (size_t)ctx->env_height;
        if (probas)
            std::memcpy(probas, ctx->env_probas_host.data(), 4 * count);
        if (alias)
            std::memcpy(alias, ctx->env_alias_host.data(), 4 * count);
        if (luminance_sum)
            *luminance_sum = ctx->env_sum;
        if (packed_texels)
            HIP_CHECK(hipMemcpy(packed_texels, ctx->env_texels.ptr, 4 * count, hipMemcpyDeviceToHost));
    });
}

int mipt_bake_settings_default(int32_t table, mipt_bake_settings* out)
{
    if (!out)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::bake_default_settings(table, out);
}

int mipt_bake_lut(mipt_ctx* c, const mipt_bake_settings* settings, float* out_device, mipt_bake_info* info)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || !settings)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] { mipt::bake_lut(*ctx, *settings, out_device, info); });
}

int mipt_set_lut(mipt_ctx* c, int32_t table, const float* values, int32_t size_cos_theta, int32_t size_roughness,
                 int32_t size_ior)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        if (table < MIPT_LUT_GGX_CONDUCTOR || table > MIPT_LUT_GGX_THIN_GLASS)
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "mipt_set_lut: the renderer reads tables 0 .. 4 only");
        mipt_bake_settings d;
        mipt::bake_default_settings(table, &d);
        if (values && (size_cos_theta != d.size_cos_theta || size_roughness != d.size_roughness ||
                       size_ior != d.size_ior))
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "mipt_set_lut: sizes differ from the renderer's table (" +
                                                      std::to_string(d.size_cos_theta) + " x " +
                                                      std::to_string(d.size_roughness) + " x " +
                                                      std::to_string(d.size_ior) + ")");
        mipt::RenderState& R = ctx->render;
        const size_t count = (size_t)d.size_cos_theta * (size_t)d.size_roughness * (size_t)d.size_ior;
        HIP_CHECK(hipDeviceSynchronize()); // queued passes may read the old table
        if (values)
            R.lut_override[table].assign(values, values + count);
        else
            R.lut_override[table].clear();
        if (R.scene_ready)
            mipt::upload_lut(R, table);
        R.graph_dirty = true;
    });
}

int mipt_get_shipped_lut(int32_t table, int32_t masking_shadowing, float* out, int32_t count)
{
    mipt_bake_settings d;
    if (!out || mipt::bake_default_settings(table, &d) != MIPT_OK || table > MIPT_LUT_GGX_THIN_GLASS)
        return MIPT_ERR_INVALID_ARGUMENT;
    if (masking_shadowing != MIPT_GGX_HEIGHT_UNCORRELATED &&
        !(masking_shadowing == MIPT_GGX_HEIGHT_CORRELATED && table == MIPT_LUT_GGX_CONDUCTOR))
        return MIPT_ERR_INVALID_ARGUMENT;
    const size_t n = (size_t)d.size_cos_theta * (size_t)d.size_roughness * (size_t)d.size_ior;
    if (count < 0 || (size_t)count != n)
        return MIPT_ERR_INVALID_ARGUMENT;
    const float* src = nullptr;
    switch (table)
    {
    case MIPT_LUT_GGX_CONDUCTOR:
        src = masking_shadowing == MIPT_GGX_HEIGHT_CORRELATED ? mipt::conductor_correlated_lut_data()
                                                               : mipt::conductor_lut_data();
        break;
    case MIPT_LUT_GLOSSY_DIELECTRIC: src = mipt::glossy_dielectric_lut_data(); break;
    case MIPT_LUT_GGX_GLASS: src = mipt::glass_lut_data(); break;
    case MIPT_LUT_GGX_GLASS_INVERSE: src = mipt::glass_inv_lut_data(); break;
    default: src = mipt::thin_glass_lut_data(); break;
    }
    std::memcpy(out, src, 4 * n);
    return MIPT_OK;
}

int mipt_compute_alias_table(const float* weights, int32_t n, float total_sum, float* out_probas, int32_t* out_alias)
{
    if (n < 0 || (n > 0 && (!weights || !out_probas || !out_alias)))
        return MIPT_ERR_INVALID_ARGUMENT;
    try
    {
        mipt::compute_alias_table(weights, n, total_sum, out_probas, out_alias);
    }
    catch (const std::bad_alloc&)
    {
        return MIPT_ERR_OUT_OF_MEMORY;
    }
    return MIPT_OK;
}

int mipt_read_restir_reservoirs(mipt_ctx* c, int32_t which, void* host_out)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || !host_out || which < 0 || which > 3)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        mipt::RenderState& R = ctx->render;
        const int32_t b = which == 3 ? R.restir_shade_buffer : which;
        if (!R.configured || !R.reservoirs[b].ptr)
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "no ReSTIR DI reservoirs: the strategy is not MIPT_LSS_RESTIR_DI");
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(host_out, R.reservoirs[b].ptr, 32 * (size_t)R.pixels, hipMemcpyDeviceToHost));
    });
}

int mipt_read_framebuffer(mipt_ctx* c, float* host_rgb, double* host_rgb64)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        if (!ctx->render.configured)
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "no framebuffer: call mipt_set_render_settings first");
        HIP_CHECK(hipDeviceSynchronize());
        const size_t n = ctx->render.pixels;
        if (host_rgb)
            HIP_CHECK(hipMemcpy(host_rgb, ctx->render.fb.ptr, 12 * n, hipMemcpyDeviceToHost));
        if (host_rgb64)
        {
            if (!ctx->render.settings.accumulate_fp64)
                mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "FP64 framebuffer requested but accumulate_fp64 is off");
            HIP_CHECK(hipMemcpy(host_rgb64, ctx->render.fb64.ptr, 24 * n, hipMemcpyDeviceToHost));
        }
    });
}

int mipt_framebuffer_device(mipt_ctx* c, float** d_rgb, double** d_rgb64)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || !ctx->render.configured)
        return MIPT_ERR_INVALID_ARGUMENT;
    if (d_rgb)
        *d_rgb = ctx->render.fb.as<float>();
    if (d_rgb64)
        *d_rgb64 = ctx->render.settings.accumulate_fp64 ? ctx->render.fb64.as<double>() : nullptr;
    return MIPT_OK;
}

int mipt_read_gbuffer_prims(mipt_ctx* c, int32_t* host_prims)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || !host_prims)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        if (!ctx->render.configured)
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "no G-buffer: call mipt_set_render_settings first");
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(host_prims, ctx->render.gb_prim.ptr, 4 * (size_t)ctx->render.pixels,
                            hipMemcpyDeviceToHost));
    });
}

int mipt_rejected_samples(mipt_ctx* c, uint32_t* out_count)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || !out_count)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        if (!ctx->render.configured)
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "no render state");
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(out_count, ctx->render.counters.ptr, 4, hipMemcpyDeviceToHost));
    });
}

void mipt_radiance_cache_settings_default(mipt_radiance_cache_settings* s)
{
    if (!s)
        return;
    std::memset(s, 0, sizeof(*s));
    s->grid_levels = 16;
    s->grid_log2_table = 19;
    s->grid_base_resolution = 16;
    s->grid_per_level_scale = 1.3819f;
    s->batch_size = 1 << 16;
    s->learning_rate = 1.0e-2f;
    s->adam_beta1 = 0.9f;
    s->adam_beta2 = 0.99f;
    s->adam_eps = 1.0e-8f;
    s->seed = 1337u;
    s->query_bounce = 1;
    s->train_every = 64;
    s->train_extension = 3;
    s->warmup_train_every = 16;
}

int mipt_set_radiance_cache(mipt_ctx* c, const mipt_radiance_cache_settings* s)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (ctx)
        ctx->render.graph_dirty = true; // a captured sample would replay stale state
    if (!ctx || !s || s->grid_levels < 1 || s->grid_levels > 16 || s->grid_log2_table < 4 ||
        s->grid_log2_table > 24 || s->grid_base_resolution < 2 || !(s->grid_per_level_scale >= 1.0f) ||
        s->batch_size < 16 || s->batch_size > (1 << 22) || !(s->learning_rate > 0.0f) ||
        !(s->adam_beta1 >= 0.0f && s->adam_beta1 < 1.0f) || !(s->adam_beta2 >= 0.0f && s->adam_beta2 < 1.0f) ||
        !(s->adam_eps >= 0.0f) || s->query_bounce < 1 || s->train_every < 1 || s->train_extension < 0 ||
        s->warmup_train_every < 1)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        if (!ctx->scene_set)
            mipt::fail(MIPT_ERR_NO_SCENE, "no scene: the radiance cache spans the scene's bounds");
        mipt::nrc_configure(*ctx, *s);
    });
}

int mipt_radiance_cache_infer(mipt_ctx* c, const float* d_queries, uint32_t n, float* d_out, void* stream)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || (n > 0 && (!d_queries || !d_out)))
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        mipt::nrc_infer(*ctx, d_queries, nullptr, n, d_out, reinterpret_cast<hipStream_t>(stream));
    });
}

int mipt_radiance_cache_encode(mipt_ctx* c, const float* d_queries, uint32_t n, uint16_t* d_out, void* stream)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || (n > 0 && (!d_queries || !d_out)))
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        mipt::nrc_encode(*ctx, d_queries, n, d_out, reinterpret_cast<hipStream_t>(stream));
    });
}

int mipt_radiance_cache_train_step(mipt_ctx* c, const float* d_queries, const float* d_targets, uint32_t n,
                                   float* out_loss, void* stream)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || (n > 0 && (!d_queries || !d_targets)))
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        if (n == 0)
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "mipt_radiance_cache_train_step: empty batch");
        mipt::nrc_gradients(*ctx, d_queries, d_targets, nullptr, n, 1u, 0u, true, s);
        mipt::nrc_apply_adam(*ctx, s);
        if (out_loss)
        {
            float sum = 0.0f;
            HIP_CHECK(hipMemcpyAsync(&sum, ctx->nrc.loss.ptr, 4, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            const uint32_t rows = std::min<uint32_t>(n, (uint32_t)ctx->nrc.settings.batch_size);
            *out_loss = sum / (3.0f * (float)rows);
        }
    });
}

int mipt_radiance_cache_gradients(mipt_ctx* c, const float* d_queries, const float* d_targets, uint32_t n,
                                  float* host_mlp_grad, float* host_grid_grad, float* out_loss)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx || n == 0 || !d_queries || !d_targets)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        mipt::Nrc& C = ctx->nrc;
        mipt::nrc_gradients(*ctx, d_queries, d_targets, nullptr, n, 1u, 0u, true, nullptr);
        HIP_CHECK(hipDeviceSynchronize());
        if (host_mlp_grad)
            HIP_CHECK(hipMemcpy(host_mlp_grad, C.w_grad.ptr, 4 * (size_t)mipt::NRC_WEIGHTS, hipMemcpyDeviceToHost));
        if (host_grid_grad)
            HIP_CHECK(hipMemcpy(host_grid_grad, C.grid_grad.ptr, 4 * C.grid_count, hipMemcpyDeviceToHost));
        if (out_loss)
        {
            float sum = 0.0f;
            HIP_CHECK(hipMemcpy(&sum, C.loss.ptr, 4, hipMemcpyDeviceToHost));
            const uint32_t rows = std::min<uint32_t>(n, (uint32_t)C.settings.batch_size);
            *out_loss = sum / (3.0f * (float)rows);
        }
        HIP_CHECK(hipMemset(C.w_grad.ptr, 0, 4 * (size_t)mipt::NRC_WEIGHTS));
        HIP_CHECK(hipMemset(C.grid_grad.ptr, 0, 4 * C.grid_count));
    });
}

int mipt_radiance_cache_get_params(mipt_ctx* c, float* mlp_master, uint16_t* mlp_bf16, float* grid_master,
                                   uint16_t* grid_f16)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        mipt::Nrc& C = ctx->nrc;
        if (!C.configured)
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "radiance cache not configured: call mipt_set_radiance_cache first");
        HIP_CHECK(hipDeviceSynchronize());
        if (mlp_master)
            HIP_CHECK(hipMemcpy(mlp_master, C.w_master.ptr, 4 * (size_t)mipt::NRC_WEIGHTS, hipMemcpyDeviceToHost));
        if (mlp_bf16)
            HIP_CHECK(hipMemcpy(mlp_bf16, C.wk.ptr, 2 * (size_t)mipt::NRC_WEIGHTS, hipMemcpyDeviceToHost));
        if (grid_master)
            HIP_CHECK(hipMemcpy(grid_master, C.grid_master.ptr, 4 * C.grid_count, hipMemcpyDeviceToHost));
        if (grid_f16)
            HIP_CHECK(hipMemcpy(grid_f16, C.grid_f16.ptr, 2 * C.grid_count, hipMemcpyDeviceToHost));
    });
}

int mipt_radiance_cache_set_params(mipt_ctx* c, const float* mlp_master, const float* grid_master)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (ctx)
        ctx->render.graph_dirty = true; // a captured sample would replay stale state
    if (!ctx)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        if (!ctx->nrc.configured)
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "radiance cache not configured: call mipt_set_radiance_cache first");
        HIP_CHECK(hipDeviceSynchronize());
        mipt::nrc_set_params(*ctx, mlp_master, grid_master);
    });
}

int32_t mipt_radiance_cache_weight_count(void) { return mipt::NRC_WEIGHTS; }

int mipt_radiance_cache_stats(mipt_ctx* c, uint32_t* counts, float* train_loss)
{
    Ctx* ctx = reinterpret_cast<Ctx*>(c);
    if (!ctx)
        return MIPT_ERR_INVALID_ARGUMENT;
    return mipt::guarded(ctx, [&] {
        mipt::Nrc& C = ctx->nrc;
        if (!C.configured || C.train_capacity == 0)
            mipt::fail(MIPT_ERR_INVALID_ARGUMENT, "no path pass with the radiance cache has run");
        HIP_CHECK(hipDeviceSynchronize());
        uint32_t h[4] = {0, 0, 0, 0};
        HIP_CHECK(hipMemcpy(h, C.counters.ptr, 12, hipMemcpyDeviceToHost));
        if (counts)
            std::memcpy(counts, h, 12);
        if (train_loss)
        {
            float sum = 0.0f;
            HIP_CHECK(hipMemcpy(&sum, C.loss.ptr, 4, hipMemcpyDeviceToHost));
            *train_loss = h[1] ? sum / (3.0f * (float)h[1]) : 0.0f;
        }
    });
}

uint32_t mipt_host_random_number(uint32_t pass_index)
{
    // Xorshift32 seeded with 42 (src/HostDeviceCommon/Xorshift.h:22-25), one
    // draw per pass (src/Renderer/GPURenderer.cpp:750-751).
    uint32_t x = 42u;
    for (uint32_t i = 0; i <= pass_index; i++)
    {
        x ^= x << 13;
        x ^= x >> 17;
        x ^= x << 5;
    }
    return x;
}

} // extern "C"
