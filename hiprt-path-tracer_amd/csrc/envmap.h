// envmap.h -- the ambient light of the scene (WorldSettings,
// src/HostDeviceCommon/WorldSettings.h:14-53): none, a uniform colour, or an
// equirectangular environment map stored as R9G9B9E5 texels and importance
// sampled by luminance through an alias table (EnvmapSamplingStrategy =
// ESS_ALIAS_TABLE, KernelOptions.h:155), with BSDF MIS
// (EnvmapSamplingDoBSDFMIS, :161) and nearest-texel lookups
// (EnvmapSamplingDoBilinearFiltering off, :170).
#pragma once

#include "internal.h"
#include "pt_math.h"
#include "shading.h"

namespace mipt {

// Device view of the world settings and the envmap's tables.
struct EnvDev {
    int32_t type;               // MIPT_AMBIENT_*
    float uniform[3];           // uniform_light_color
    int32_t scale_background;   // envmap_scale_background_intensity
    float intensity;            // envmap_intensity
    float total_sum;            // luminance sum of the float texels (before packing)
    int32_t width, height;
    const uint32_t* texels;     // R9G9B9E5, row y at v = y / height
    const float* probas;        // alias table over the texels
    const int32_t* alias;
    float to_world[3][3];       // envmap_to_world_matrix, [column][row]
    float to_env[3][3];         // world_to_envmap_matrix
};

EnvDev make_env_dev(const Ctx& ctx);

__device__ __forceinline__ bool env_sampled(const EnvDev& e)
{
    return e.type == MIPT_AMBIENT_ENVMAP && e.intensity > 0.0f && e.width > 0;
}

// matrix_X_vec (Math.h:636-649): column-major, summed left to right.
__device__ __forceinline__ pt::f3 env_rotate(const float m[3][3], pt::f3 v)
{
    return pt::mk(m[0][0] * v.x + m[1][0] * v.y + m[2][0] * v.z, m[0][1] * v.x + m[1][1] * v.y + m[2][1] * v.z,
                  m[0][2] * v.x + m[1][2] * v.y + m[2][2] * v.z);
}

// sample_environment_map_texture (Texture.h:209-240), nearest texel; the
// texel coordinates are clamped to the image (u, v are in [0, 1] for any
// finite direction).
__device__ __forceinline__ pt::Color env_texel(const EnvDev& e, float u, float v)
{
    const int32_t x = min(max((int32_t)(u * (float)(uint32_t)(e.width - 1)), 0), e.width - 1);
    const int32_t y = min(max((int32_t)(v * (float)(uint32_t)(e.height - 1)), 0), e.height - 1);
    float r, g, b;
    pt::rgb9e5_unpack(e.texels[x + y * e.width], r, g, b);
    return pt::col(r, g, b) * e.intensity;
}

// eval_envmap_no_pdf (Envmap.h:29-38).
__device__ __forceinline__ pt::Color env_radiance(const EnvDev& e, pt::f3 world_dir)
{
    const pt::f3 d = env_rotate(e.to_env, world_dir);
    const float u = 0.5f + pt::atan2_fixed(d.z, d.x) * 0x1.45f306p-3f; // 1 / (2 pi)
    const float v = 0.5f + pt::asin_fixed(d.y) * 0x1.45f306p-2f;      // 1 / pi
    return env_texel(e, u, v);
}

// The luminance pdf of a texel's radiance in solid angle at polar angle
// theta (Envmap.h:122-131, :155-163).
__device__ __forceinline__ float env_pdf_of(const EnvDev& e, const pt::Color& radiance, float sin_theta)
{
    float pdf = pt::luminance(radiance) / (e.total_sum * e.intensity);
    pdf = pdf * (float)((uint32_t)e.width * (uint32_t)e.height);
    return pdf / (0x1.3bd3ccp+4f * sin_theta); // 2 pi^2
}

// envmap_sample (Envmap.h:77-133): a texel by the alias table, its corner's
// direction (Y up), the radiance looked up at that direction's (u, v).
__device__ __forceinline__ pt::Color env_sample(const EnvDev& e, pt::f3& dir, float& pdf, pt::Xorshift32& rng)
{
    int32_t i = rng.random_index(e.width * e.height);
    if (rng() > e.probas[i])
        i = e.alias[i];
    const int32_t y = (int32_t)floorf((float)i / (float)(uint32_t)e.width);
    const int32_t x = (int32_t)floorf((float)i - (float)y * (float)(uint32_t)e.width);
    const float u = (float)x / (float)(uint32_t)e.width;
    const float v = (float)y / (float)(uint32_t)e.height;
    const float phi = u * 0x1.921fb6p+2f;                            // 2 pi
    const float theta = fmaxf(1.0e-5f, v * 0x1.921fb6p+1f);          // pi
    float sin_t, cos_t, sin_p, cos_p;
    pt::sincos_fixed(theta, sin_t, cos_t);
    pt::sincos_fixed(phi, sin_p, cos_p);
    dir = env_rotate(e.to_world, pt::mk(-sin_t * cos_p, -cos_t, -sin_t * sin_p));
    const pt::Color radiance = env_texel(e, u, v);
    pdf = env_pdf_of(e, radiance, sin_t);
    return radiance;
}

// envmap_eval (Envmap.h:139-166): radiance and pdf of a world direction; the
// polar angle is taken from the world direction, before the rotation.
__device__ __forceinline__ pt::Color env_eval(const EnvDev& e, pt::f3 dir, float& pdf)
{
    const pt::Color radiance = env_radiance(e, dir);
    const float theta = pt::acos_fixed(-dir.y);
    pdf = env_pdf_of(e, radiance, pt::sin_fixed(theta));
    return radiance;
}

// path_tracing_miss_gather_envmap (PathTracing.h:121-170) without clamps
// (the contribution clamps default to 0, off): what a ray that leaves the
// scene at bounce `bounce` adds, before the throughput.
__device__ __forceinline__ pt::Color env_miss(const EnvDev& e, pt::f3 dir, int32_t bounce)
{
    if (e.type == MIPT_AMBIENT_UNIFORM)
        return pt::col(e.uniform[0], e.uniform[1], e.uniform[2]);
    if (e.type != MIPT_AMBIENT_ENVMAP || e.intensity == 0.0f || bounce != 0)
        return pt::col(0.0f);
    // the envmap is importance sampled: only camera rays see it
    pt::Color c = env_radiance(e, dir);
    if (!e.scale_background)
        c = c / e.intensity;
    return c;
}

} // namespace mipt
