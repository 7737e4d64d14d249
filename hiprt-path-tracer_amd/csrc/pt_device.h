// pt_device.h -- the per-pass kernel argument (the HIPRTRenderData subset the
// passes read) and the surface / light-sampling device functions shared by the
// path-tracing and ReSTIR DI kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "envmap.h"
#include "internal.h"
#include "shading.h"

namespace mipt {

using pt::Color;
using pt::f3;

// Passed by value to every pass kernel, like HIPRTRenderData
// (src/HostDeviceCommon/RenderData.h:93-129).
struct PtData {
    // scene
    const float* verts;
    const int32_t* idx;
    const int32_t* mat_idx;
    const pt::Material* mats;
    const int32_t* mat_class; // lobe class of each material (the shading order's sort key)
    // batched views: first pixel of each view (null with one view)
    const ViewDev* views;
    int32_t view_count;
    // graph replay: per-sample values in device memory (null: the fields below)
    const FrameParams* frame;
    const int32_t* emissive;
    const float* areas;
    // smooth shading (get_shading_normal): per-vertex flag and normals, null
    // when the scene has none or do_normal_mapping is 0
    const uint8_t* has_normals;
    const float* vnormals;
    int32_t emissive_count;
    // light sampling (LSS_BASE_UNIFORM / LSS_BASE_POWER)
    int32_t light_base;
    const float* alias_probas;
    const int32_t* alias_alias;
    float alias_sum;
    // camera
    float cam_pos[3];
    // settings
    int32_t width, height;
    int32_t nb_bounces;
    int32_t do_rr, rr_min_depth;
    float rr_clamp;
    int32_t freeze, enable_direct;
    int32_t scene_volumes;     // some material transmits light (glass or diffuse transmission)
    int32_t scene_dispersive;  // some material is dispersive glass
    uint32_t sample_number;
    uint32_t random_number;
    uint32_t train_only_every; // > 0: a cache warm-up pass traces only its training paths
    pt::BsdfGlobals g;
    // G-buffer
    int32_t* gb_prim;
    float4* gb_pos;
    uint32_t* gb_gn;
    uint32_t* gb_sn;
    int32_t* gb_mat;
    // path state
    float4* throughput;   // xyz
    float4* radiance;     // xyz = ray_color
    float4* vtx_pos;      // xyz = inter point, w = accumulated roughness
    float4* vtx_sn;       // xyz = shading normal, w = int bits of prim
    float4* vtx_gn;       // xyz = geometric normal, w = int bits of material
    float4* vtx_view;     // xyz = view direction (towards the viewer)
    float4* nee_light;    // xyz = pending light-sample contribution
    float4* nee_bsdf;     // xyz = bsdf_color * cos, w = bsdf pdf
    uint32_t* rng;
    uint32_t* flags;
    // volume state of the paths whose state is not canonical (FLAG_VOL):
    // vol_a = stack entries 1-3 + wavelength bits, vol_b = pos | inside << 8,
    // incident, outgoing, distance bits. Entry 0 is always the air.
    uint4* vol_a;
    uint4* vol_b;
    // rays of the current bounce
    float4* shadow_rays;
    uint8_t* shadow_occ;
    float4* bsdf_rays;
    int4* bsdf_hits;
    float4* cont_rays;
    int4* cont_hits;
    // ambient light; with envmap sampling, the pending contributions of the
    // envmap light sample and of its MIS BSDF sample (shadow slots n + pix,
    // 2n + pix)
    EnvDev env;
    float4* env_light;
    float4* env_bsdf;
    // RIS of the lights (sample_lights_RIS, RIS/RIS.h:252): at every bounce
    // (LSS_RIS_BSDF_AND_LIGHT) or after bounce 0 under ReSTIR DI; candidate
    // counts; per path, the light candidate kept and the reservoir's state
    // until the BSDF candidate's ray is back (3 float4)
    int32_t ris_all, ris_later;
    int32_t ris_light, ris_bsdf;
    float4* ris_state;
    // ReSTIR DI at bounce 0 (LSS_RESTIR_DI): the reservoirs to shade
    int32_t restir;
    int32_t final_shading_visibility;
    const float4* restir_out; // 2 float4 per pixel
    // output
    float* fb;        // rgb sum
    double* fb64;     // rgb sum (may be null)
    uint32_t* rejected;
};

PtData make_pt_data(Ctx& ctx, uint32_t sample_number, uint32_t random_number);

// The sample's values: from the graph's device record when replaying.
__device__ __forceinline__ uint32_t frame_sample(const PtData& d)
{
    return d.frame ? d.frame->sample_number : d.sample_number;
}
__device__ __forceinline__ uint32_t frame_random(const PtData& d)
{
    return d.frame ? d.frame->rn_path : d.random_number;
}
__device__ __forceinline__ pt::BsdfGlobals frame_globals(const PtData& d)
{
    pt::BsdfGlobals g = d.g;
    if (d.frame)
    {
        g.consistent_tau = d.frame->consistent_tau;
        g.consistent_tau_refract = d.frame->consistent_tau_refract;
    }
    return g;
}

// View of pixel `pix` of a batch (last view whose first pixel is <= pix).
__device__ __forceinline__ int32_t view_of(const ViewDev* views, int32_t count, uint32_t pix)
{
    int32_t lo = 0, hi = count - 1;
    while (lo < hi)
    {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (views[mid].start <= pix)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// The pixel's index within its own view: what the per-pixel random seed of
// a single-view render uses (Megakernel.h:52-57).
__device__ __forceinline__ uint32_t view_pixel(const PtData& d, uint32_t pix)
{
    if (!d.views)
        return pix;
    return pix - d.views[view_of(d.views, d.view_count, pix)].start;
}

__device__ __forceinline__ f3 ld3(float4 v) { return pt::mk(v.x, v.y, v.z); }
__device__ __forceinline__ float4 st3(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
__device__ __forceinline__ Color ldc(float4 v) { return pt::col(v.x, v.y, v.z); }
__device__ __forceinline__ float4 stc(Color c, float w) { return make_float4(c.r, c.g, c.b, w); }

__device__ __forceinline__ void tri_verts(const PtData& d, int32_t prim, f3& a, f3& b, f3& c)
{
    const int32_t* t = d.idx + 3 * (int64_t)prim;
    const float* v0 = d.verts + 3 * (int64_t)t[0];
    const float* v1 = d.verts + 3 * (int64_t)t[1];
    const float* v2 = d.verts + 3 * (int64_t)t[2];
    a = pt::mk(v0[0], v0[1], v0[2]);
    b = pt::mk(v1[0], v1[1], v1[2]);
    c = pt::mk(v2[0], v2[1], v2[2]);
}

// get_triangle_normal_not_normalized (LightUtils.h:19-31)
__device__ __forceinline__ f3 tri_normal_unnormalized(const PtData& d, int32_t prim)
{
    f3 a, b, c;
    tri_verts(d, prim, a, b, c);
    return pt::cross(b - a, c - a);
}

// Ray slots: float4 (origin, last_prim bits), float4 (direction, tmax).
__device__ __forceinline__ void write_inactive(float4* rays, size_t i)
{
    rays[2 * i] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
    rays[2 * i + 1] = make_float4(0.0f, 0.0f, 1.0f, -1.0f);
}

__device__ __forceinline__ void write_ray(float4* rays, size_t i, f3 o, int32_t last, f3 dir, float tmax)
{
    rays[2 * i] = make_float4(o.x, o.y, o.z, __int_as_float(last));
    rays[2 * i + 1] = make_float4(dir.x, dir.y, dir.z, tmax);
}

// Intersect.h:142-176
__device__ __forceinline__ void fix_backfacing(f3 view, f3& gn, f3& sn)
{
    if (pt::dot(view, gn) < 0.0f)
    {
        gn = -1.0f * gn;
        sn = -1.0f * sn;
    }
    if (pt::dot(view, sn) < 0.0f)
        sn = -1.0f * sn;
    f3 refl = pt::reflect_ray(view, sn);
    if (pt::dot(refl, gn) <= 0.0f)
    {
        const float eps = 0.01f;
        refl = refl - pt::normalize((pt::dot(refl, gn) - eps) * gn);
        sn = pt::normalize(view + refl);
    }
}

// Shading normal (get_shading_normal, Intersect.h:114-133): when the
// triangle's first vertex has a vertex normal, the normalised barycentric
// interpolation n_B u + n_C v + n_A (1 - u - v) (uv_interpolate,
// Texture.h:248-251), else the geometric normal.
__device__ __forceinline__ f3 shading_normal(const PtData& d, int32_t prim, float u, float v, f3 gn)
{
    if (!d.has_normals)
        return gn;
    const int32_t* t = d.idx + 3 * (int64_t)prim;
    if (!d.has_normals[t[0]])
        return gn;
    const float* na = d.vnormals + 3 * (int64_t)t[0];
    const float* nb = d.vnormals + 3 * (int64_t)t[1];
    const float* nc = d.vnormals + 3 * (int64_t)t[2];
    const float w = (1.0f - u) - v;
    const f3 n = (u * pt::mk(nb[0], nb[1], nb[2]) + v * pt::mk(nc[0], nc[1], nc[2])) + w * pt::mk(na[0], na[1], na[2]);
    return pt::normalize(n);
}

// Hit attributes of trace_main_path_ray (Intersect.h:222-245): geometric
// normal, shading normal, back-face fix.
__device__ __forceinline__ void hit_attributes(const PtData& d, f3 o, f3 dir, int32_t prim, float t, float u, float v,
                                               f3& p, f3& gn, f3& sn)
{
    f3 a, b, c;
    tri_verts(d, prim, a, b, c);
    p = o + t * dir;
    gn = pt::normalize(pt::normalize(pt::cross(b - a, c - a)));
    sn = shading_normal(d, prim, u, v, gn);
    fix_backfacing(-1.0f * dir, gn, sn);
}

// Per-path volume state in memory (see PtData::vol_a / vol_b).
__device__ __forceinline__ void vol_load(const PtData& d, uint32_t pix, pt::VolumeState& v)
{
    const uint4 a = d.vol_a[pix], b = d.vol_b[pix];
    v.e0 = pt::vol_entry(pt::VOL_AIR, 0u, true, true);
    v.e1 = a.x;
    v.e2 = a.y;
    v.e3 = a.z;
    v.wavelength = __uint_as_float(a.w);
    v.pos = (int32_t)(b.x & 0xffu);
    v.inside = (int32_t)(b.x >> 8);
    v.incident = (int32_t)b.y;
    v.outgoing = (int32_t)b.z;
    v.distance = __uint_as_float(b.w);
    v.reset = 0;
}

__device__ __forceinline__ void vol_store(const PtData& d, uint32_t pix, const pt::VolumeState& v)
{
    d.vol_a[pix] = make_uint4(v.e1, v.e2, v.e3, __float_as_uint(v.wavelength));
    d.vol_b[pix] = make_uint4((uint32_t)v.pos | ((uint32_t)v.inside << 8), (uint32_t)v.incident, (uint32_t)v.outgoing,
                              __float_as_uint(v.distance));
}

// The state right after the first push of `mat` on an empty stack: what a
// path without a stored state has at a vertex of that material.
__device__ __forceinline__ void vol_canonical_at(pt::VolumeState& v, int32_t mat, int32_t priority)
{
    pt::vol_init(v);
    pt::vol_push(v, mat, priority);
}

// Dispersive glass: hitting it samples the path's wavelength
// (Intersect.h:264-272, RayVolumeState.h:50-58).
__device__ __forceinline__ bool dispersive(const pt::Material& m)
{
    return m.dispersion_scale > 0.0f && m.specular_transmission > 0.0f;
}

// get_dispersion_ray_color (Dispersion.h:512-535): the colour filter of the
// sampled wavelength, applied once (the wavelength turns positive).
__device__ __forceinline__ Color dispersion_filter(float& wavelength, float dispersion_scale)
{
    if (dispersion_scale == 0.0f)
        return pt::col(1.0f);
    if (wavelength >= 0.0f)
        return pt::col(1.0f);
    wavelength *= -1.0f;
    const f3 c = pt::wavelength_to_rgb(wavelength);
    return pt::col(c.x, c.y, c.z);
}

// Octahedral24BitNormal::unpack (Packing.h:237-360): decode then normalize.
__device__ __forceinline__ f3 gb_normal(uint32_t p) { return pt::normalize(pt::oct_unpack(p)); }

// ColorRGB32F::luminance (Color.h:90)
__device__ __forceinline__ float luminance(Color c) { return 0.3086f * c.r + 0.6094f * c.g + 0.0820f * c.b; }

// compute_cosine_term_at_light_source (PDFConversion.h:19-30), back faces rejected.
__device__ __forceinline__ float cos_at_light(f3 n, f3 minus_dir) { return pt::maxf_(0.0f, pt::dot(n, minus_dir)); }

// area_to_solid_angle_pdf (PDFConversion.h:32-38)
__device__ __forceinline__ float area_to_solid_angle(float area_pdf, float dist, float cos_l)
{
    if (cos_l < 1.0e-8f)
        return 0.0f;
    return area_pdf * (dist * dist) / cos_l;
}

// balance_heuristic (Sampling.h:103-116)
__device__ __forceinline__ float balance_heuristic(float pa, float na, float pb, float nb)
{
    if (pa == 0.0f)
        return 0.0f;
    return pa / (na * pa + nb * pb);
}

// LightSampleInformation (src/HostDeviceCommon/LightSampleInformation.h:14-30)
struct LightSample {
    int32_t tri;
    f3 normal;
    float area;
    Color emission;
    f3 point;
    float area_pdf;
};

// sample_one_emissive_triangle_uniform / _power (LightUtils.h:248-309) with
// sample_point_on_generic_triangle (:163-199, Heitz 2019 square-to-triangle
// :142-156) and DeviceAliasTable::sample (AliasTable.h:13-22). Requires
// emissive_count > 0.
__device__ __forceinline__ LightSample sample_emissive_with(const PtData& d, int32_t strategy, pt::Xorshift32& rng)
{
    int32_t e = rng.random_index(d.emissive_count);
    if (strategy == MIPT_LSS_BASE_POWER)
    {
        const float p = d.alias_probas[e];
        if (rng() > p)
            e = d.alias_alias[e];
    }
    LightSample s;
    s.tri = d.emissive[e];
    f3 A, B, C;
    tri_verts(d, s.tri, A, B, C);
    float r1 = rng();
    float r2 = rng();
    if (r2 > r1)
    {
        r1 *= 0.5f;
        r2 -= r1;
    }
    else
    {
        r2 *= 0.5f;
        r1 -= r2;
    }
    const f3 AB = B - A, AC = C - A;
    const f3 nrm = pt::cross(AB, AC);
    const float len = sqrtf(pt::dot(nrm, nrm));
    s.point = (A + r1 * AB) + r2 * AC;
    s.normal = pt::vdiv(nrm, len);
    s.area = 0.5f * len;
    s.emission = d.mats[d.mat_idx[s.tri]].emission;
    s.area_pdf = 1.0f / s.area;
    if (strategy == MIPT_LSS_BASE_POWER)
        s.area_pdf *= (luminance(s.emission) * s.area) / d.alias_sum;
    else
        s.area_pdf /= (float)d.emissive_count;
    return s;
}

// With the render settings' base strategy.
__device__ __forceinline__ LightSample sample_emissive(const PtData& d, pt::Xorshift32& rng)
{
    return sample_emissive_with(d, d.light_base, rng);
}

// pdf_of_emissive_triangle_hit_area_measure (LightUtils.h:51-76)
__device__ __forceinline__ float emissive_hit_area_pdf_with(const PtData& d, int32_t strategy, float light_area,
                                                            Color emission)
{
    float pdf = 1.0f / light_area;
    if (strategy == MIPT_LSS_BASE_POWER)
        pdf *= (luminance(emission) * light_area) / d.alias_sum;
    else
        pdf /= (float)d.emissive_count;
    return pdf;
}

__device__ __forceinline__ float emissive_hit_area_pdf(const PtData& d, float light_area, Color emission)
{
    return emissive_hit_area_pdf_with(d, d.light_base, light_area, emission);
}

} // namespace mipt
