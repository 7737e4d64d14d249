// envmap.hip -- GPURenderer::set_envmap (GPURenderer.cpp:962-998) on the
// MI355X: the float image goes to HBM once and one thread per texel packs it
// to R9G9B9E5 (RGBE9995Envmap::pack_from, EnvmapRGBE9995.h:21-44; 4 bytes a
// texel, 32 MB for 4096 x 2048). The alias table over the texels' luminance
// and the luminance sum follow Image32Bit::compute_alias_table
// (Image.cpp:652-672): a sequential binary32 sum in row order, which fixes
// every pdf the kernels compute, so it is built on the host like the
// emissive power table.
#include <cstring>
#include <vector>

#include "envmap.h"
#include "internal.h"

namespace mipt {

namespace {

__global__ void k_pack_envmap(const float* __restrict__ pixels, int32_t channels, uint32_t count,
                              uint32_t* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    const float* p = pixels + (size_t)i * (size_t)channels;
    out[i] = pt::rgb9e5_pack(p[0], p[1], p[2]);
}

} // namespace

bool env_sampling_on(const Ctx& ctx)
{
    return ctx.world.ambient_light_type == MIPT_AMBIENT_ENVMAP && ctx.world.envmap_intensity > 0.0f &&
           ctx.env_width > 0;
}

// Shadow ray slots of the path pass: the light sample's, and with envmap
// sampling the envmap sample's and the envmap MIS BSDF sample's.
void allocate_env_path_buffers(Ctx& ctx)
{
    RenderState& R = ctx.render;
    const size_t n = R.pixels;
    if (n == 0)
        return; // no render buffers yet: mipt_set_render_settings allocates them
    const uint32_t slots = env_sampling_on(ctx) ? 3u : 1u;
    if (slots > R.shadow_slots || !R.shadow_rays.ptr)
    {
        R.shadow_rays.reserve(32 * n * slots);
        R.shadow_occ.reserve(n * slots);
        R.q_shadow.reserve(4 * (size_t)SEG_COUNT * seg_capacity((uint32_t)n) * slots);
        R.shadow_slots = slots;
    }
    if (slots == 3u)
    {
        R.env_light.reserve(16 * n);
        R.env_bsdf.reserve(16 * n);
    }
}

void set_envmap(Ctx& ctx, const float* pixels, int32_t width, int32_t height, int32_t channels)
{
    if (width == 0 || height == 0)
    {
        // an empty envmap: back to the uniform ambient light
        if (ctx.world.ambient_light_type == MIPT_AMBIENT_ENVMAP)
            ctx.world.ambient_light_type = MIPT_AMBIENT_UNIFORM;
        ctx.env_width = ctx.env_height = 0;
        return;
    }
    if (width < 0 || height < 0 || !pixels || (channels != 3 && channels != 4))
        fail(MIPT_ERR_INVALID_ARGUMENT, "mipt_set_envmap: pixels of width x height texels of 3 or 4 floats");
    const uint64_t count64 = (uint64_t)width * (uint64_t)height;
    if (count64 > (1ull << 26))
        fail(MIPT_ERR_INVALID_ARGUMENT, "mipt_set_envmap: at most 2^26 texels (8192 x 8192)");
    const uint32_t count = (uint32_t)count64;
    // luminance_of_pixel (Image.cpp:526-546) summed in row order
    std::vector<float> lum(count);
    float sum = 0.0f;
    for (uint32_t i = 0; i < count; i++)
    {
        const float* p = pixels + (size_t)i * (size_t)channels;
        float l = 0.0f;
        l += p[0] * 0.3086f;
        l += p[1] * 0.6094f;
        l += p[2] * 0.0820f;
        lum[i] = l;
        sum += l;
    }
    std::vector<float> probas(count);
    std::vector<int32_t> alias(count);
    compute_alias_table(lum.data(), (int32_t)count, sum, probas.data(), alias.data());

    DeviceBuffer staging;
    staging.reserve(sizeof(float) * (size_t)count * (size_t)channels);
    HIP_CHECK(hipMemcpy(staging.ptr, pixels, sizeof(float) * (size_t)count * (size_t)channels,
                        hipMemcpyHostToDevice));
    ctx.env_texels.reserve(4 * (size_t)count);
    hipLaunchKernelGGL(k_pack_envmap, dim3((count + 255) / 256), dim3(256), 0, nullptr, staging.as<float>(), channels,
                       count, ctx.env_texels.as<uint32_t>());
    HIP_CHECK(hipGetLastError());
    ctx.env_probas.reserve(4 * (size_t)count);
    ctx.env_alias.reserve(4 * (size_t)count);
    HIP_CHECK(hipMemcpy(ctx.env_probas.ptr, probas.data(), 4 * (size_t)count, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(ctx.env_alias.ptr, alias.data(), 4 * (size_t)count, hipMemcpyHostToDevice));
    HIP_CHECK(hipDeviceSynchronize());
    ctx.env_probas_host = std::move(probas);
    ctx.env_alias_host = std::move(alias);
    ctx.env_sum = sum;
    ctx.env_width = width;
    ctx.env_height = height;
    ctx.world.ambient_light_type = MIPT_AMBIENT_ENVMAP;
    allocate_env_path_buffers(ctx);
}

EnvDev make_env_dev(const Ctx& ctx)
{
    EnvDev e{};
    const mipt_world_settings& w = ctx.world;
    e.type = w.ambient_light_type;
    if (e.type == MIPT_AMBIENT_ENVMAP && ctx.env_width == 0)
        e.type = MIPT_AMBIENT_NONE; // unreachable: mipt_set_world_settings refuses it
    for (int k = 0; k < 3; k++)
        e.uniform[k] = w.uniform_light_color[k];
    e.scale_background = w.envmap_scale_background_intensity;
    e.intensity = w.envmap_intensity;
    e.total_sum = ctx.env_sum;
    e.width = ctx.env_width;
    e.height = ctx.env_height;
    e.texels = ctx.env_texels.as<uint32_t>();
    e.probas = ctx.env_probas.as<float>();
    e.alias = ctx.env_alias.as<int32_t>();
    std::memcpy(e.to_world, w.envmap_to_world, sizeof(e.to_world));
    std::memcpy(e.to_env, w.world_to_envmap, sizeof(e.to_env));
    return e;
}

} // namespace mipt
