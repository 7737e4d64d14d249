// pathtrace.hip -- the per-pass launch surface of the reference renderer on
// the wavefront design:
//
//   CAMERA_RAYS  (CameraRays, src/Device/kernels/CameraRays.h:83-213)
//       [reset at sample 0] -> k_camera_rays -> closest-hit traversal -> k_gbuffer_write
//   PATH_TRACE   (MegaKernel, src/Device/kernels/Megakernel.h:23-144)
//       k_path_init, then per bounce:
//         k_shade   : direct light (NEE light sample + MIS BSDF sample, or the
//                     ReSTIR DI reservoir at bounce 0), continuation sample and
//                     Russian roulette (all RNG draws, in the reference's order)
//         traversal : any hit (shadow rays), closest hit (BSDF-to-light rays,
//                     continuation rays)
//         k_resolve : combines the ray results into the path radiance
//                     (Lights.h:123-226, :364-423) and builds the next vertex
//                     (Intersect.h:222-245)
//       k_accumulate: framebuffer sum (PathTracing.h:176-217) + FP64 mirror
//
// With the learned radiance cache on, k_cache_query runs at the start of the
// query bounce: a path that is not a training path stops there and records a
// query; after the bounces the cache is evaluated on the queries and
// k_cache_resolve adds throughput x prediction to their pixels. One path in
// train_every (a hash of pixel and sample) is a training path: it records the
// same query and its radiance so far, runs to the bounce limit, and
// k_cache_targets turns its remaining radiance / throughput into the target.
//
// Each path carries its xorshift32 state between kernels, so the per-path
// RNG sequence is the megakernel's. Rays of a bounce live in fixed per-path
// slots; an inactive slot has tmax = -1 and leaves the traversal at the root.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "nrc.h"
#include "pt_device.h"

namespace mipt {
namespace {

constexpr uint32_t FLAG_ALIVE = 1u;      // path is at a surface vertex
constexpr uint32_t FLAG_NEE_LIGHT = 2u;  // shadow ray pending with a contribution
constexpr uint32_t FLAG_NEE_BSDF = 4u;   // BSDF-to-light ray pending
constexpr uint32_t FLAG_CONTINUE = 8u;   // continuation ray traced this bounce
constexpr uint32_t FLAG_DIRECT = 16u;    // direct lighting estimated this bounce
constexpr uint32_t FLAG_VOL = 32u;       // volume state stored in vol_a / vol_b
constexpr uint32_t FLAG_SKIP = 64u;      // continuation ray goes on past a skipped boundary
// The light sample's BSDF evaluation floors the distance in the medium at
// 1e-6 (Beer-Lambert); the reference evaluates it only when the shadow ray is
// unoccluded, so the floor is applied after the shadow ray, at the next vertex.
constexpr uint32_t FLAG_NEE_FLOOR = 128u;
// Envmap sampling (Envmap.h:168-255): the envmap light sample's shadow ray
// (slot n + pix) and the MIS BSDF sample's (slot 2n + pix) are pending; the
// light sample's evaluation floored the medium distance (as FLAG_NEE_FLOOR).
constexpr uint32_t FLAG_ENV_LIGHT = 256u;
constexpr uint32_t FLAG_ENV_BSDF = 512u;
constexpr uint32_t FLAG_ENV_FLOOR = 1024u;
// Direct light by RIS this bounce: the light candidates' reservoir is in
// ris_state, its kept light sample's shadow ray under FLAG_NEE_LIGHT and the
// BSDF candidate's ray under FLAG_NEE_BSDF.
constexpr uint32_t FLAG_RIS = 2048u;

// Slots of the live rays of one bounce, per ray kind (segmented lists), and
// the two lists the boundary-skip rounds alternate between.
struct RayQueues {
    SegList shadow;
    SegList bsdf;
    SegList cont;
    SegList skip[2];
};

// Zeroes the counters of `lists` segmented lists stored back to back.
__global__ void k_seg_reset(uint32_t* counts, uint32_t lists)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < lists * SEG_COUNT)
        counts[i * SEG_STRIDE] = 0u;
}

// ---------------------------------------------------------------------------
// CAMERA_RAYS
// ---------------------------------------------------------------------------
struct GBufferView {
    int32_t* prim;
    float4* pos;
    uint32_t* gn;
    uint32_t* sn;
    int32_t* mat;
};

// reset_render (CameraRays.h:20-59): the G-buffers go back to "no hit". The
// reset normal bits are 0 and the reset material index -1; neither is read
// before a hit overwrites it.
__global__ void k_gbuffer_reset(GBufferView g, uint32_t n)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= n)
        return;
    g.prim[pix] = -1;
    g.gn[pix] = 0u;
    g.sn[pix] = 0u;
    g.pos[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    g.mat[pix] = -1;
}

// CameraRays.h:117-124 (previous-frame copy) and :176-201 (G-buffer write
// after the closest hit).
__global__ void k_gbuffer_write(PtData d, const float4* __restrict__ rays, const int4* __restrict__ hits,
                                GBufferView prev, int32_t copy_prev)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= (uint32_t)(d.width * d.height))
        return;
    if (copy_prev)
    {
        prev.gn[pix] = d.gb_gn[pix];
        prev.sn[pix] = d.gb_sn[pix];
        prev.mat[pix] = d.gb_mat[pix];
        prev.pos[pix] = d.gb_pos[pix];
        prev.prim[pix] = d.gb_prim[pix];
    }
    const float4 ra = rays[2 * (size_t)pix], rb = rays[2 * (size_t)pix + 1];
    const f3 o = pt::mk(ra.x, ra.y, ra.z), dir = pt::mk(rb.x, rb.y, rb.z);
    const int4 h = hits[pix];
    if (h.x >= 0)
    {
        f3 p, gn, sn;
        hit_attributes(d, o, dir, h.x, __int_as_float(h.y), __int_as_float(h.z), __int_as_float(h.w), p, gn, sn);
        d.gb_gn[pix] = pt::oct_pack(gn);
        d.gb_sn[pix] = pt::oct_pack(sn);
        d.gb_mat[pix] = d.mat_idx[h.x];
        d.gb_pos[pix] = st3(p, 0.0f);
    }
    else
        d.gb_pos[pix] = st3(o + dir, 0.0f);
    d.gb_prim[pix] = h.x >= 0 ? h.x : -1;
}

// ---------------------------------------------------------------------------
// PATH_TRACE
// ---------------------------------------------------------------------------

// One path in `every` of a sample is a training path of the radiance cache
// (a hash of pixel and sample number).
__device__ __forceinline__ bool is_training_path(uint32_t pix, uint32_t sample, uint32_t every)
{
    uint32_t h = pix * 0x9E3779B1u ^ (sample + 0x7F4A7C15u) * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return h % every == 0u;
}

// Megakernel.h:52-86: seed, first vertex from the G-buffer.
__device__ __forceinline__ bool init_path(const PtData& d, uint32_t pix)
{
    if (d.train_only_every && !is_training_path(pix, frame_sample(d), d.train_only_every))
    {
        d.flags[pix] = 0u; // warm-up pass: only the training paths run
        return false;
    }
    d.rng[pix] = pt::pixel_seed(view_pixel(d, pix), frame_sample(d), frame_random(d), d.freeze != 0);
    d.throughput[pix] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    d.radiance[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const int32_t prim = d.gb_prim[pix];
    const f3 p = ld3(d.gb_pos[pix]);
    f3 cam = pt::mk(d.cam_pos[0], d.cam_pos[1], d.cam_pos[2]);
    if (d.views)
    {
        const ViewDev& v = d.views[view_of(d.views, d.view_count, pix)];
        cam = pt::mk(v.position[0], v.position[1], v.position[2]);
    }
    // GBufferDevice.h:19-22: view = normalize(camera - hit); the path ray is
    // normalize(-view) and the BSDF view direction is -ray.direction. A pixel
    // whose camera ray missed holds origin + direction as its hit point.
    const f3 ray_dir = pt::normalize(-1.0f * pt::normalize(cam - p));
    if (prim < 0)
    {
        // path_tracing_miss_gather_envmap at bounce 0 (PathTracing.h:121-170)
        const Color bg = env_miss(d.env, ray_dir, 0) * pt::col(1.0f);
        d.radiance[pix] = stc(bg, 0.0f);
        d.flags[pix] = 0u;
        return false;
    }
    const f3 gn = gb_normal(d.gb_gn[pix]);
    const f3 sn = gb_normal(d.gb_sn[pix]);
    d.vtx_pos[pix] = st3(p, 0.0f);
    d.vtx_sn[pix] = st3(sn, __int_as_float(prim));
    d.vtx_gn[pix] = st3(gn, __int_as_float(d.gb_mat[pix]));
    d.vtx_view[pix] = st3(-1.0f * ray_dir, 0.0f);
    uint32_t flags = FLAG_ALIVE;
    // RayVolumeState::reconstruct_first_hit (RayVolumeState.h:35-59): the
    // camera hit's material is pushed on an empty stack (implicit: a path
    // without FLAG_VOL has exactly that state); a dispersive glass samples
    // the wavelength with the path's first draw.
    if (d.scene_dispersive)
    {
        const int32_t mi = d.mat_idx[prim];
        const pt::Material& m = d.mats[mi];
        if (dispersive(m))
        {
            pt::VolumeState vs;
            vol_canonical_at(vs, mi, m.dielectric_priority);
            pt::Xorshift32 rng{d.rng[pix]};
            vs.wavelength = -pt::wavelength_from_draw(rng());
            d.rng[pix] = rng.state;
            vol_store(d, pix, vs);
            flags |= FLAG_VOL;
        }
    }
    d.flags[pix] = flags;
    return true;
}

// Starts every pixel's path and lists the live ones: the later kernels of
// the pass run over that list, not over the image.
__global__ void k_path_init(PtData d, SegList live)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    const bool alive = pix < (uint32_t)(d.width * d.height) && init_path(d, pix);
    seg_append(live, alive, pix);
}

// Pixel of dense index w of a list of live paths; false past its end. Every
// lane of the wave must call it; a wave past the end leaves at once.
struct LiveIndex {
    uint32_t pix;
    bool valid;
    bool wave_empty;
};

__device__ __forceinline__ LiveIndex live_index(const SegList& l)
{
    SegMap map;
    map.init(l.counts, l.seg_cap);
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t total = map.total;
    LiveIndex r{0u, false, (w & ~63u) >= total};
    if (r.wave_empty)
        return r;
    const size_t slot = map.slot(w < total ? w : total - 1u);
    if (w < total)
    {
        r.pix = l.slots[slot];
        r.valid = true;
    }
    return r;
}

// evaluate_ReSTIR_DI_reservoir (ReSTIR/DI/FinalShading.h:16-101) without the
// shadow ray: returns the contribution and whether a shadow ray must decide it.
__device__ __forceinline__ Color restir_final_shading(const PtData& d, uint32_t pix, const pt::Material& m, f3 p, f3 sn,
                                                      f3 gn, f3 view, pt::VolumeState& vs, bool& need_shadow,
                                                      f3& dir_out, float& dist_out)
{
    need_shadow = false;
    const float4 ra = d.restir_out[2 * (size_t)pix], rb = d.restir_out[2 * (size_t)pix + 1];
    const float ucw = rb.z;
    const int32_t tri = __float_as_int(rb.y);
    const uint32_t flags = (uint32_t)__float_as_int(rb.w);
    const bool env = (flags & pt::RESTIR_FLAG_ENVMAP) != 0u;
    // A NaN weight sum leaves UCW = NaN with no sample selected (triangle -1,
    // no envmap flag): nothing to shade. validate_reservoir (:103-108): an
    // envmap sample without the envmap contributes nothing.
    if (ucw <= 0.0f || (tri < 0 && !env) || (env && !env_sampled(d.env)))
        return pt::col(0.0f);
    float dist;
    f3 dir;
    if (env)
    {
        dir = env_rotate(d.env.to_world, ld3(ra));
        dist = 1.0e35f;
    }
    else
    {
        const f3 to = ld3(ra) - p;
        dist = sqrtf(pt::dot(to, to));
        dir = pt::vdiv(to, dist);
    }
    dir_out = dir;
    dist_out = dist;
    if (!(flags & pt::RESTIR_FLAG_UNOCCLUDED) && d.final_shading_visibility)
        need_shadow = true;
    pt::BsdfCtx c{view, sn, gn, dir, flags & pt::RESTIR_FLAG_LOBE_MASK, pt::REGULARIZATION_MIS, 0.0f, &vs, 0};
    float bpdf;
    const Color f = pt::principled_eval(frame_globals(d), m, c, bpdf);
    float cosine = pt::dot(sn, dir);
    if (flags & pt::FROM_GLASS_REFRACT)
        cosine = fabsf(cosine);
    if (!(cosine > 0.0f))
        return pt::col(0.0f);
    Color emission;
    float conv = 1.0f;
    if (env)
    {
        float epdf;
        emission = env_eval(d.env, dir, epdf);
    }
    else
    {
        emission = d.mats[d.mat_idx[tri]].emission;
        conv = cos_at_light(pt::normalize(tri_normal_unnormalized(d, tri)), -1.0f * dir);
        conv /= dist * dist;
    }
    return (((f * ucw) * emission) * cosine) * conv;
}

// Lights.h:123-226 (NEE with MIS), Lights.h:261-294 (ReSTIR DI at bounce 0)
// and PathTracing.h:81-101 (next bounce + Russian roulette).
__device__ __forceinline__ void shade_vertex(const PtData& d, int32_t bounce, uint32_t pix, uint32_t flags,
                                             bool& shadow_active, bool& bsdf_active, bool& cont, bool& env_light_active,
                                             bool& env_bsdf_active)
{
    const bool stored_volume = (flags & FLAG_VOL) != 0u;
    flags &= FLAG_ALIVE;
    const float4 vp = d.vtx_pos[pix], vsn = d.vtx_sn[pix], vg = d.vtx_gn[pix];
    const f3 p = ld3(vp), sn = ld3(vsn), gn = ld3(vg), view = ld3(d.vtx_view[pix]);
    const int32_t prim = __float_as_int(vsn.w);
    const int32_t mat_index = __float_as_int(vg.w);
    const pt::Material& m = d.mats[mat_index];
    float path_roughness = vp.w;
    pt::Xorshift32 rng{d.rng[pix]};
    bool light_floor = false;
    // the path's volume state after the push of this vertex's material
    pt::VolumeState vs;
    if (stored_volume)
        vol_load(d, pix, vs);
    else
        vol_canonical_at(vs, mat_index, m.dielectric_priority);

    if (bounce > 0 || d.enable_direct)
    {
        flags |= FLAG_DIRECT;
        d.nee_light[pix] = stc(pt::col(0.0f), 0.0f);
    }
    // Without emissive triangles sample_emissive_geometry returns before any
    // draw (Lights.h:313-321): neither the light nor the MIS BSDF sample runs.
    // ReSTIR DI at bounce 0 also covers the envmap (Lights.h:318-324)
    if ((bounce > 0 || d.enable_direct) && (d.emissive_count > 0 || env_sampled(d.env)) && d.restir && bounce == 0)
    {
        bool need_shadow;
        f3 dir;
        float dist;
        const Color contrib = restir_final_shading(d, pix, m, p, sn, gn, view, vs, need_shadow, dir, dist);
        d.nee_light[pix] = stc(contrib, 0.0f);
        flags |= FLAG_NEE_LIGHT;
        if (need_shadow)
        {
            write_ray(d.shadow_rays, pix, p, prim, dir, dist - 1.0e-4f);
            shadow_active = true;
        }
    }
    else if ((bounce > 0 || d.enable_direct) && d.emissive_count > 0 && (d.ris_all || (d.ris_later && bounce > 0)))
    {
        // ---- sample_bsdf_and_lights_RIS_reservoir (RIS/RIS.h:83-244): the
        // light candidates stream through the reservoir here; the BSDF
        // candidate's weight needs its ray's hit and joins in resolve_path.
        // can_do_light_sampling is true with microfacet regularisation
        // (MaterialUtils.h:71-78). The candidates' evaluations update the
        // path's volume state, as the reference's do.
        const float nl = (float)d.ris_light, nb = (float)d.ris_bsdf;
        float wsum = 0.0f, target_l = 0.0f, cos_l = 0.0f, dist_l = 0.0f;
        bool has_light = false;
        Color f_l = pt::col(0.0f), em_l = pt::col(0.0f);
        f3 dir_l = pt::mk(0.0f, 0.0f, 0.0f);
        for (int32_t i = 0; i < d.ris_light; i++)
        {
            const LightSample ls = sample_emissive(d, rng);
            float w = 0.0f, target = 0.0f, cos_p = 0.0f, dist = 0.0f;
            Color f = pt::col(0.0f);
            f3 dirn = pt::mk(0.0f, 0.0f, 0.0f);
            if (ls.area_pdf > 0.0f)
            {
                const f3 to = ls.point - p;
                dist = sqrtf(pt::dot(to, to));
                dirn = pt::vdiv(to, dist);
                const float cos_at_light = pt::maxf_(0.0f, pt::dot(ls.normal, -1.0f * dirn));
                cos_p = fabsf(pt::dot(sn, dirn));
                if (cos_p > 0.0f && cos_at_light > 1.0e-6f)
                {
                    pt::BsdfCtx c{view, sn, gn, dirn, pt::NO_INFO, pt::REGULARIZATION_MIS, path_roughness, &vs, 0};
                    float bsdf_pdf = 0.0f;
                    f = pt::principled_eval(frame_globals(d), m, c, bsdf_pdf);
                    target = pt::luminance(f * ls.emission * cos_p);
                    const float spdf =
                        cos_at_light < 1.0e-8f ? 0.0f : ls.area_pdf * (dist * dist) / cos_at_light;
                    w = balance_heuristic(spdf, nl, bsdf_pdf, nb) * target / spdf;
                }
            }
            wsum += w;
            if (rng() < w / wsum)
            {
                has_light = true;
                f_l = f;
                em_l = ls.emission;
                target_l = target;
                cos_l = cos_p;
                dir_l = dirn;
                dist_l = dist;
            }
        }
        float u_b = 0.0f;
        if (d.ris_bsdf > 0)
        {
            pt::BsdfCtx c{view, sn, gn, pt::mk(0.0f, 0.0f, 0.0f), pt::NO_INFO, pt::REGULARIZATION_MIS, path_roughness,
                          &vs, 0};
            f3 dir;
            float bpdf;
            const Color f = pt::principled_sample(frame_globals(d), m, c, dir, bpdf, rng);
            if (bpdf > 0.0f)
            {
                d.nee_bsdf[pix] = stc(f, bpdf);
                write_ray(d.bsdf_rays, pix, p, prim, dir, 1.0e35f - 1.0e-4f);
                bsdf_active = true;
                flags |= FLAG_NEE_BSDF;
            }
            u_b = rng();
        }
        if (has_light)
        {
            write_ray(d.shadow_rays, pix, p, prim, dir_l, dist_l - 1.0e-4f);
            shadow_active = true;
            flags |= FLAG_NEE_LIGHT;
        }
        float4* rs = d.ris_state + 3 * (size_t)pix;
        rs[0] = stc(f_l, cos_l);
        rs[1] = stc(em_l, target_l);
        rs[2] = make_float4(wsum, u_b, has_light ? 1.0f : 0.0f, 0.0f);
        flags |= FLAG_RIS;
    }
    else if ((bounce > 0 || d.enable_direct) && d.emissive_count > 0)
    {
        // ---- light sample (LightUtils.h:248-309, :142-199)
        Color light_mis = pt::col(0.0f);
        {
            const LightSample ls = sample_emissive(d, rng);
            if (ls.area_pdf > 0.0f)
            {
                const f3 to = ls.point - p;
                const float dist = sqrtf(pt::dot(to, to));
                const f3 dirn = pt::vdiv(to, dist);
                pt::VolumeState vs_light = vs;
                pt::BsdfCtx c{view, sn, gn, dirn, pt::NO_INFO, pt::REGULARIZATION_MIS, path_roughness, &vs_light, 0};
                float bsdf_pdf;
                const Color f = pt::principled_eval(frame_globals(d), m, c, bsdf_pdf);
                if (vs_light.distance != vs.distance)
                {
                    light_floor = true;
                    write_ray(d.shadow_rays, pix, p, prim, dirn, dist - 1.0e-4f);
                    shadow_active = true;
                }
                if (bsdf_pdf > 0.0f)
                {
                    const float cos_l = pt::maxf_(0.0f, pt::dot(ls.normal, -1.0f * dirn));
                    if (cos_l > 1.0e-5f)
                    {
                        float lpdf = cos_l < 1.0e-8f ? 0.0f : ls.area_pdf * (dist * dist) / cos_l;
                        const float w = lpdf == 0.0f ? 0.0f : lpdf / (1.0f * lpdf + 1.0f * bsdf_pdf);
                        const float cosine = fabsf(pt::dot(sn, dirn));
                        light_mis = f * cosine * ls.emission * w / lpdf / 1.0f;
                        write_ray(d.shadow_rays, pix, p, prim, dirn, dist - 1.0e-4f);
                        shadow_active = true;
                        flags |= FLAG_NEE_LIGHT;
                    }
                }
            }
        }
        d.nee_light[pix] = stc(light_mis, 0.0f);
        // ---- BSDF sample for MIS (Lights.h:174-223)
        pt::BsdfCtx c{view, sn, gn, pt::mk(0.0f, 0.0f, 0.0f), pt::NO_INFO, pt::REGULARIZATION_MIS, path_roughness, &vs,
                      0};
        f3 dir;
        float bpdf;
        const Color f = pt::principled_sample(frame_globals(d), m, c, dir, bpdf, rng);
        if (bpdf > 0.0f)
        {
            const float cosine = fabsf(pt::dot(sn, dir));
            d.nee_bsdf[pix] = stc(f * cosine, bpdf);
            write_ray(d.bsdf_rays, pix, p, prim, dir, 1.0e35f - 1.0e-4f);
            bsdf_active = true;
            flags |= FLAG_NEE_BSDF;
        }
    }
    // sample_environment_map (Envmap.h:168-276): after the emissive
    // geometry, and at bounce 0 only without ReSTIR DI (which would sample
    // the envmap among its candidates)
    bool env_floor = false;
    if ((bounce > 0 || d.enable_direct) && env_sampled(d.env) && !(bounce == 0 && d.restir))
    {
        const uint32_t n = (uint32_t)(d.width * d.height);
        f3 edir;
        float epdf;
        const Color ecol = env_sample(d.env, edir, epdf, rng);
        // can_do_light_sampling is true with microfacet regularisation
        // (MaterialUtils.h:71-78)
        const float ecos = pt::dot(sn, edir);
        Color env_light = pt::col(0.0f);
        if (epdf > 0.0f && ecos > 0.0f)
        {
            pt::VolumeState vs_env = vs;
            pt::BsdfCtx c{view, sn, gn, edir, pt::NO_INFO, pt::REGULARIZATION_MIS, path_roughness, &vs_env, 0};
            float bsdf_pdf;
            const Color f = pt::principled_eval(frame_globals(d), m, c, bsdf_pdf);
            const float w = balance_heuristic(epdf, 1.0f, bsdf_pdf, 1.0f);
            env_light = f * ecos * w * ecol / epdf / 1.0f;
            env_floor = vs_env.distance != vs.distance;
            write_ray(d.shadow_rays, n + pix, p, prim, edir, 1.0e35f - 1.0e-4f);
            env_light_active = true;
            flags |= FLAG_ENV_LIGHT;
        }
        d.env_light[pix] = stc(env_light, 0.0f);
        pt::BsdfCtx c{view, sn, gn, pt::mk(0.0f, 0.0f, 0.0f), pt::NO_INFO, pt::REGULARIZATION_MIS, path_roughness, &vs,
                      0};
        f3 bdir;
        float bpdf;
        const Color f = pt::principled_sample(frame_globals(d), m, c, bdir, bpdf, rng);
        const float bcos = fabsf(pt::dot(sn, bdir));
        if (bpdf > 0.0f)
        {
            // envmap_eval needs no ray: only its visibility is pending
            float env_pdf;
            const Color radiance = env_eval(d.env, bdir, env_pdf);
            if (env_pdf > 0.0f)
            {
                const float w = balance_heuristic(bpdf, 1.0f, env_pdf, 1.0f);
                d.env_bsdf[pix] = stc(radiance * w * bcos * f / bpdf, 0.0f);
                write_ray(d.shadow_rays, 2 * n + pix, p, prim, bdir, 1.0e35f - 1.0e-4f);
                env_bsdf_active = true;
                flags |= FLAG_ENV_BSDF;
            }
        }
    }
    if (!shadow_active)
        d.shadow_occ[pix] = 0; // a contribution without a shadow ray is unoccluded

    // ---- continuation (PathTracing.h:26-101)
    {
        pt::BsdfCtx c{view, sn, gn, pt::mk(0.0f, 0.0f, 0.0f), pt::NO_INFO, pt::NO_REGULARIZATION, path_roughness, &vs,
                      1};
        f3 dir;
        float bpdf;
        const Color f = pt::principled_sample(frame_globals(d), m, c, dir, bpdf, rng);
        // RayPayload::accumulate_roughness (RayPayload.h:52-89)
        if (c.info == pt::FROM_DIFFUSE || c.info == pt::FROM_DIFFUSE_TRANSMISSION)
            path_roughness = 1.0f;
        else if (c.info == pt::FROM_COAT)
            path_roughness = pt::maxf_(m.coat_roughness, path_roughness);
        else if (c.info == pt::FROM_FIRST_METAL || c.info == pt::FROM_SPECULAR || c.info == pt::FROM_GLASS_REFLECT ||
                 c.info == pt::FROM_GLASS_REFRACT)
            path_roughness = pt::maxf_(m.roughness, path_roughness);
        else if (c.info == pt::FROM_SECOND_METAL)
            path_roughness = pt::maxf_(m.second_roughness, path_roughness);
        if (bpdf > 0.0f)
        {
            Color thr = ldc(d.throughput[pix]);
            const Color att = f * fabsf(pt::dot(dir, sn)) / bpdf;
            bool survive = true;
            if (bounce >= d.rr_min_depth && d.do_rr)
            {
                float sp = pt::max_component(thr);
                sp = pt::minf_(sp, 1.0f);
                if (rng() > sp)
                    survive = false;
                else
                {
                    float scale = 1.0f / sp;
                    if (d.rr_clamp > 0.0f)
                        scale = pt::minf_(scale, d.rr_clamp);
                    thr = thr * scale;
                }
            }
            if (survive)
            {
                thr = thr * dispersion_filter(vs.wavelength, m.dispersion_scale);
                thr = thr * att;
                thr = pt::col(pt::maxf_(thr.r, 1.0e-5f), pt::maxf_(thr.g, 1.0e-5f), pt::maxf_(thr.b, 1.0e-5f));
                if (!pt::is_black(thr) && bounce < d.nb_bounces)
                {
                    d.throughput[pix] = stc(thr, 0.0f);
                    write_ray(d.cont_rays, pix, p, prim, dir, INFINITY);
                    cont = true;
                    flags |= FLAG_CONTINUE;
                }
            }
        }
    }
    // The direct-light estimate of this vertex uses the pre-update throughput,
    // which k_shade keeps in thr_prev before calling this.
    d.vtx_pos[pix] = st3(p, path_roughness);
    d.rng[pix] = rng.state;
    // A reflection pops the vertex's material again: most paths are back to
    // the canonical state and keep nothing in memory.
    if (light_floor && !vs.reset)
        flags |= FLAG_NEE_FLOOR;
    if (env_floor && !vs.reset)
        flags |= FLAG_ENV_FLOOR;
    if (!pt::vol_canonical(vs) || (flags & (FLAG_NEE_FLOOR | FLAG_ENV_FLOOR)))
    {
        vol_store(d, pix, vs);
        flags |= FLAG_VOL;
    }
    d.flags[pix] = flags;
}

// ---------------------------------------------------------------------------
// Material-sorted shading order. The live list of a bounce is sorted by the
// lobe class of the material at each path's vertex (and whether the path
// carries a volume state) with a counting sort: k_sort_hist counts the
// classes (a block's histogram in LDS, then one atomic per class present on
// sub-counter blockIdx % 16), k_sort_scatter writes each path to
// start(class, sub) + the block's base (one atomic per class present) + its
// rank inside the block, waves and lanes in list order. A class's paths from
// one block stay contiguous and in pixel order, so the per-pixel state reads
// of k_shade stay mostly coalesced.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t shade_class(const PtData& d, uint32_t pix)
{
    const uint32_t flags = d.flags[pix];
    if (!(flags & FLAG_ALIVE))
        return 0u;
    const int32_t mi = __float_as_int(d.vtx_gn[pix].w);
    return (uint32_t)d.mat_class[mi] | ((flags & FLAG_VOL) ? 32u : 0u);
}

__global__ void k_sort_reset(uint32_t* counts)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2u * SHADE_SORT_COUNTERS)
        counts[i * SEG_STRIDE] = 0u;
}

__global__ __launch_bounds__(256) void k_sort_hist(PtData d, SegList live, uint32_t* __restrict__ counts)
{
    __shared__ uint32_t h[SHADE_SORT_CLASSES];
    if (threadIdx.x < SHADE_SORT_CLASSES)
        h[threadIdx.x] = 0u;
    __syncthreads();
    const LiveIndex li = live_index(live);
    if (li.valid)
        atomicAdd(&h[shade_class(d, li.pix)], 1u);
    __syncthreads();
    const uint32_t c = threadIdx.x;
    if (c < SHADE_SORT_CLASSES && h[c] > 0u)
        atomicAdd(&counts[(c * SHADE_SORT_SUBS + blockIdx.x % SHADE_SORT_SUBS) * SEG_STRIDE], h[c]);
}

__global__ __launch_bounds__(256) void k_sort_scatter(PtData d, SegList live, const uint32_t* __restrict__ counts,
                                                      uint32_t* __restrict__ cursors, uint32_t* __restrict__ out)
{
    constexpr int WAVES = 256 / 64;
    __shared__ uint32_t wave_count[WAVES][SHADE_SORT_CLASSES];
    __shared__ uint32_t wave_base[WAVES][SHADE_SORT_CLASSES];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t sub = blockIdx.x % SHADE_SORT_SUBS;
    wave_count[wave][lane] = 0u;
    __syncthreads();
    const LiveIndex li = live_index(live);
    const bool valid = li.valid;
    const uint32_t cls = valid ? shade_class(d, li.pix) : 0u;
    // rank among the wave's lanes of the same class, one round per class
    // present in the wave
    uint32_t rank = 0u;
    uint64_t todo = __ballot(valid);
    while (todo)
    {
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)todo) - 1u;
        const uint32_t c = __shfl(cls, (int)leader, 64);
        const uint64_t same = __ballot(valid && cls == c);
        if (valid && cls == c)
            rank = (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        if (lane == leader)
            wave_count[wave][c] = (uint32_t)__popcll(same);
        todo &= ~same;
    }
    __syncthreads();
    if (threadIdx.x < SHADE_SORT_CLASSES)
    {
        // start of (class c, sub-counter sub): every path of the classes
        // before c, then those of c on the sub-counters before sub
        const uint32_t c = threadIdx.x;
        uint32_t total_c = 0u, before_sub = 0u;
        for (uint32_t s2 = 0; s2 < SHADE_SORT_SUBS; s2++)
        {
            const uint32_t v = counts[(c * SHADE_SORT_SUBS + s2) * SEG_STRIDE];
            total_c += v;
            if (s2 < sub)
                before_sub += v;
        }
        uint32_t incl = total_c;
        for (int off = 1; off < 64; off <<= 1)
        {
            const uint32_t t = __shfl_up(incl, off, 64);
            if (lane >= (uint32_t)off)
                incl += t;
        }
        uint32_t block_total = 0u;
        for (int w = 0; w < WAVES; w++)
            block_total += wave_count[w][c];
        uint32_t base = incl - total_c + before_sub;
        if (block_total > 0u)
            base += atomicAdd(&cursors[(c * SHADE_SORT_SUBS + sub) * SEG_STRIDE], block_total);
        for (int w = 0; w < WAVES; w++)
        {
            wave_base[w][c] = base;
            base += wave_count[w][c];
        }
    }
    __syncthreads();
    if (valid)
        out[wave_base[wave][cls] + rank] = li.pix;
}

// Shades the live paths of the list (dense, so that no wave idles on dead
// or stopped paths) and enqueues the slots of the rays they emitted; keeps
// the throughput before the vertex for the direct-light estimate. With
// `sorted`, the list's paths are taken in that order (k_sort_scatter).
// The layered BSDF with every lobe needs ~140 VGPRs: allowing them (3 waves per
// SIMD) beats spilling to scratch at 4 waves (11.08 vs 11.74 ms per 1080p
// config-1 sample).
__global__ __launch_bounds__(256, 2) void k_shade(PtData d, int32_t bounce, RayQueues q, SegList live,
                                                  const uint32_t* __restrict__ sorted, float4* __restrict__ thr_prev)
{
    const LiveIndex li = live_index(live);
    if (li.wave_empty)
        return;
    const uint32_t pix = sorted && li.valid ? sorted[blockIdx.x * blockDim.x + threadIdx.x] : li.pix;
    const uint32_t flags = li.valid ? d.flags[pix] : 0u;
    bool shadow_active = false, bsdf_active = false, cont = false, env_light = false, env_bsdf = false;
    if (flags & FLAG_ALIVE)
    {
        thr_prev[pix] = d.throughput[pix];
        shade_vertex(d, bounce, pix, flags, shadow_active, bsdf_active, cont, env_light, env_bsdf);
    }
    const uint32_t n = (uint32_t)(d.width * d.height);
    seg_append(q.shadow, shadow_active, pix);
    seg_append(q.shadow, env_light, n + pix);
    seg_append(q.shadow, env_bsdf, 2 * n + pix);
    seg_append(q.bsdf, bsdf_active, pix);
    seg_append(q.cont, cont, pix);
}

// The continuation ray's closest hit (trace_main_path_ray, Intersect.h:200-275):
// the boundary's material is pushed on the interior stack (paths without a
// stored state rebuild theirs at the vertex); a boundary of lower priority
// than the medium the ray is in is skipped and the ray goes on from the hit
// point with the same self-intersection filter (returns true). Hitting a
// dispersive glass samples the path's wavelength.
enum : int { PATH_ENDED = 0, PATH_AT_VERTEX = 1, PATH_SKIPPING = 2 };

__device__ __forceinline__ int next_vertex(const PtData& d, uint32_t pix, uint32_t flags)
{
    const int4 h = d.cont_hits[pix];
    const float4 ra = d.cont_rays[2 * (size_t)pix], rb = d.cont_rays[2 * (size_t)pix + 1];
    const f3 o = pt::mk(ra.x, ra.y, ra.z), dir = pt::mk(rb.x, rb.y, rb.z);
    if (h.x < 0)
    {
        // path_tracing_miss_gather_envmap (PathTracing.h:121-170) at a bounce
        // > 0: the uniform ambient colour (a sampled envmap adds nothing)
        if (d.env.type == MIPT_AMBIENT_UNIFORM)
        {
            const Color add = env_miss(d.env, dir, 1) * ldc(d.throughput[pix]);
            d.radiance[pix] = stc(ldc(d.radiance[pix]) + add, 0.0f);
        }
        d.flags[pix] = 0u;
        return PATH_ENDED;
    }
    const float t = __int_as_float(h.y);
    f3 p, gn, sn;
    hit_attributes(d, o, dir, h.x, t, __int_as_float(h.z), __int_as_float(h.w), p, gn, sn);
    const int32_t mi = d.mat_idx[h.x];
    bool stored = (flags & FLAG_VOL) != 0u;
    pt::VolumeState vs;
    if (stored)
    {
        vol_load(d, pix, vs);
        if ((flags & FLAG_NEE_FLOOR) && !d.shadow_occ[pix])
            vs.distance = pt::maxf_(vs.distance, 1.0e-6f);
        if ((flags & FLAG_ENV_FLOOR) && !d.shadow_occ[(size_t)(d.width * d.height) + pix])
            vs.distance = pt::maxf_(vs.distance, 1.0e-6f);
        const bool skip = pt::vol_push(vs, mi, d.mats[mi].dielectric_priority);
        if (vs.inside)
            vs.distance += t;
        if (skip)
        {
            vs.distance += t;
            vol_store(d, pix, vs);
            write_ray(d.cont_rays, pix, p, __float_as_int(ra.w), dir, INFINITY);
            d.flags[pix] = FLAG_VOL | FLAG_SKIP;
            return PATH_SKIPPING;
        }
    }
    if (d.scene_dispersive)
    {
        const pt::Material& m = d.mats[mi];
        if (!stored)
            vol_canonical_at(vs, mi, m.dielectric_priority);
        if (dispersive(m) && vs.wavelength == 0.0f)
        {
            pt::Xorshift32 rng{d.rng[pix]};
            vs.wavelength = -pt::wavelength_from_draw(rng());
            d.rng[pix] = rng.state;
            stored = true;
        }
    }
    if (stored)
        vol_store(d, pix, vs);
    d.vtx_pos[pix] = st3(p, d.vtx_pos[pix].w);
    d.vtx_sn[pix] = st3(sn, __int_as_float(h.x));
    d.vtx_gn[pix] = st3(gn, __int_as_float(mi));
    d.vtx_view[pix] = st3(-1.0f * dir, 0.0f);
    d.flags[pix] = FLAG_ALIVE | (stored ? FLAG_VOL : 0u);
    return PATH_AT_VERTEX;
}

// Direct lighting of the bounce (Lights.h:364-398) and the next vertex;
// returns the path's state after the continuation ray (PATH_*).
// The BSDF candidate joins the RIS reservoir (RIS/RIS.h:196-241), then the
// kept sample is shaded (RIS.h:18-80): a light sample through its shadow ray,
// a BSDF sample with the contribution and cosine of its candidate.
__device__ __forceinline__ Color resolve_ris(const PtData& d, uint32_t pix, uint32_t flags)
{
    const float4* rs = d.ris_state + 3 * (size_t)pix;
    const float4 s2 = rs[2];
    const float nl = (float)d.ris_light, nb = (float)d.ris_bsdf;
    float wsum = s2.x;
    bool bsdf_kept = false;
    Color f_b = pt::col(0.0f), em_b = pt::col(0.0f);
    float cos_b = 0.0f, target_b = 0.0f;
    if (d.ris_bsdf > 0)
    {
        float w = 0.0f;
        bool hit_light = false;
        if (flags & FLAG_NEE_BSDF)
        {
            const int4 h = d.bsdf_hits[pix];
            if (h.x >= 0)
            {
                const int32_t hp = h.x;
                const Color emission = d.mats[d.mat_idx[hp]].emission;
                const float4 rb = d.bsdf_rays[2 * (size_t)pix + 1];
                const f3 dir = pt::mk(rb.x, rb.y, rb.z);
                f3 a, b, c;
                tri_verts(d, hp, a, b, c);
                const f3 hgn = pt::normalize(pt::normalize(pt::cross(b - a, c - a)));
                const float cos_at_light = pt::maxf_(0.0f, pt::dot(hgn, -1.0f * dir));
                if (!pt::is_black(emission) && cos_at_light > 0.0f)
                {
                    const float4 nbv = d.nee_bsdf[pix];
                    const float bpdf = nbv.w;
                    f_b = ldc(nbv);
                    em_b = emission;
                    cos_b = fabsf(pt::dot(ld3(d.vtx_sn[pix]), dir));
                    target_b = pt::luminance(f_b * emission * cos_b);
                    const float pdf_area = emissive_hit_area_pdf(d, d.areas[hp], emission);
                    const float t = __int_as_float(h.y);
                    const float lpdf = cos_at_light < 1.0e-8f ? 0.0f : pdf_area * (t * t) / cos_at_light;
                    w = balance_heuristic(bpdf, nb, lpdf, nl) * target_b / bpdf;
                    hit_light = true;
                }
            }
        }
        wsum += w;
        if (s2.y < w / wsum)
            bsdf_kept = hit_light;
    }
    const float4 s0 = rs[0], s1 = rs[1];
    const float target = bsdf_kept ? target_b : s1.w;
    const float ucw = wsum == 0.0f ? 0.0f : 1.0f / target * wsum;
    if (!(ucw > 0.0f))
        return pt::col(0.0f);
    if (bsdf_kept)
        return f_b * ucw * em_b * cos_b;
    if (s2.z == 0.0f || d.shadow_occ[pix])
        return pt::col(0.0f);
    return ldc(s0) * ucw * ldc(s1) * s0.w;
}

__device__ __forceinline__ int resolve_path(const PtData& d, int32_t bounce, const float4* __restrict__ thr_prev,
                                            uint32_t pix)
{
    const uint32_t flags = d.flags[pix];
    if (!(flags & FLAG_ALIVE))
        return PATH_ENDED;
    Color radiance = ldc(d.radiance[pix]);
    if (flags & FLAG_DIRECT)
    {
        // estimate_direct_lighting (Lights.h:364-398)
        Color light_mis = pt::col(0.0f);
        if ((flags & FLAG_NEE_LIGHT) && !d.shadow_occ[pix])
            light_mis = ldc(d.nee_light[pix]);
        Color bsdf_mis = pt::col(0.0f);
        if (flags & FLAG_RIS)
            light_mis = resolve_ris(d, pix, flags);
        else if (flags & FLAG_NEE_BSDF)
        {
            const int4 h = d.bsdf_hits[pix];
            if (h.x >= 0)
            {
                const int32_t hp = h.x;
                const Color emission = d.mats[d.mat_idx[hp]].emission;
                const float4 rb = d.bsdf_rays[2 * (size_t)pix + 1];
                const f3 dir = pt::mk(rb.x, rb.y, rb.z);
                f3 a, b, c;
                tri_verts(d, hp, a, b, c);
                const f3 hgn = pt::normalize(pt::normalize(pt::cross(b - a, c - a)));
                const float cos_l = pt::maxf_(0.0f, pt::dot(hgn, -1.0f * dir));
                if (!pt::is_black(emission) && cos_l > 0.0f)
                {
                    const float pdf_area = emissive_hit_area_pdf(d, d.areas[hp], emission);
                    const float t = __int_as_float(h.y);
                    const float lpdf = cos_l < 1.0e-8f ? 0.0f : pdf_area * (t * t) / cos_l;
                    const float4 nb = d.nee_bsdf[pix];
                    const float bpdf = nb.w;
                    const float w = bpdf == 0.0f ? 0.0f : bpdf / (1.0f * bpdf + 1.0f * lpdf);
                    bsdf_mis = ldc(nb) * emission * w / bpdf;
                }
            }
        }
        const Color direct = light_mis + bsdf_mis;
        // the envmap's BSDF-sampled part, then its light-sampled part
        // (Envmap.h:253)
        const size_t n = (size_t)(d.width * d.height);
        Color env_bsdf = pt::col(0.0f), env_light = pt::col(0.0f);
        if ((flags & FLAG_ENV_BSDF) && !d.shadow_occ[2 * n + pix])
            env_bsdf = ldc(d.env_bsdf[pix]);
        if ((flags & FLAG_ENV_LIGHT) && !d.shadow_occ[n + pix])
            env_light = ldc(d.env_light[pix]);
        Color total = pt::col(0.0f);
        if (bounce == 0)
            total = total + d.mats[__float_as_int(d.vtx_gn[pix].w)].emission;
        total = total + ((direct / 1.0f) + (env_bsdf + env_light)) * ldc(thr_prev[pix]);
        radiance = radiance + total;
        d.radiance[pix] = stc(radiance, 0.0f);
    }
    if (!(flags & FLAG_CONTINUE))
    {
        d.flags[pix] = 0u;
        return PATH_ENDED;
    }
    return next_vertex(d, pix, flags);
}

// Over the paths shaded this bounce: those at a new vertex go to `next`, those
// whose continuation ray skipped a boundary to `skip`.
__global__ void k_resolve(PtData d, int32_t bounce, const float4* __restrict__ thr_prev, SegList live, SegList skip,
                          SegList next)
{
    const LiveIndex li = live_index(live);
    if (li.wave_empty)
        return;
    const int st = li.valid ? resolve_path(d, bounce, thr_prev, li.pix) : PATH_ENDED;
    seg_append(skip, st == PATH_SKIPPING, li.pix);
    seg_append(next, st == PATH_AT_VERTEX, li.pix);
}

// One round of boundary skipping: the paths of `in` have traced their ray on
// from the skipped boundary; those that skip again go to `out`, those at a
// vertex to `next`. The grid covers every pixel; waves past the list's length
// leave at once.
__global__ void k_resolve_skip(PtData d, SegList in, SegList out, SegList next)
{
    const LiveIndex li = live_index(in);
    if (li.wave_empty)
        return;
    const int st = li.valid ? next_vertex(d, li.pix, d.flags[li.pix]) : PATH_ENDED;
    seg_append(out, st == PATH_SKIPPING, li.pix);
    seg_append(next, st == PATH_AT_VERTEX, li.pix);
}

// SanityCheck.h:78-106 + PathTracing.h:176-217: a NaN/inf/negative sample is
// counted and dropped (contributes 0 at sample 0), otherwise summed.
__global__ void k_accumulate(PtData d)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= (uint32_t)(d.width * d.height))
        return;
    float4 r = d.radiance[pix];
    bool bad = false;
    const float v[3] = {r.x, r.y, r.z};
    for (int k = 0; k < 3; k++)
        bad |= v[k] < 0.0f || isnan(v[k]) || isinf(v[k]);
    if (bad)
    {
        atomicAdd(d.rejected, 1u);
        r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (frame_sample(d) != 0)
            return;
    }
    float* fb = d.fb + 3 * (size_t)pix;
    if (frame_sample(d) == 0)
    {
        fb[0] = r.x;
        fb[1] = r.y;
        fb[2] = r.z;
    }
    else
    {
        fb[0] += r.x;
        fb[1] += r.y;
        fb[2] += r.z;
    }
    if (d.fb64)
    {
        double* f64 = d.fb64 + 3 * (size_t)pix;
        if (frame_sample(d) == 0)
        {
            f64[0] = r.x;
            f64[1] = r.y;
            f64[2] = r.z;
        }
        else
        {
            f64[0] += (double)r.x;
            f64[1] += (double)r.y;
            f64[2] += (double)r.z;
        }
    }
}

// ---------------------------------------------------------------------------
// Learned radiance cache records
// ---------------------------------------------------------------------------
struct CacheView {
    float4* queries;          // 4 float4 per query: cache input + throughput in floats 13..15
    uint32_t* query_pix;
    const float4* query_out;  // cache prediction per query
    float4* train_queries;    // same layout as queries
    uint32_t* train_pix;
    float4* train_rad;        // path radiance before the query vertex
    float4* train_targets;
    uint32_t* counters;       // [0] queries, [1] training records, [2] targets dropped as non-finite
    SegList query_list;       // pixels of the queries, as the query kernel appends them
    SegList train_list;       // pixels of the training records
    uint32_t train_every;
    uint32_t train_capacity;
    uint32_t sample_number;
};

__global__ void k_cache_counters_reset(uint32_t* c)
{
    if (threadIdx.x < 4)
        c[threadIdx.x] = 0u;
}

// At the start of the query bounce: lists the live paths as queries (their
// path stops here) or training records. The training list is the bounce's
// list of live paths from here on.
__global__ void k_cache_query(PtData d, CacheView cv, SegList live)
{
    const LiveIndex li = live_index(live);
    if (li.wave_empty)
        return;
    const uint32_t pix = li.pix;
    const uint32_t flags = li.valid ? d.flags[pix] : 0u;
    const bool alive = (flags & FLAG_ALIVE) != 0u;
    const bool train = alive && is_training_path(pix, cv.sample_number, cv.train_every);
    seg_append(cv.query_list, alive && !train, pix);
    seg_append(cv.train_list, train, pix);
    if (alive && !train)
        d.flags[pix] = 0u; // the cache stands in for the rest of the path
}

// The vertex's cache input (position, view direction, shading normal,
// roughness, base colour) and the path throughput, as a 16-float record.
__device__ __forceinline__ void write_cache_record(const PtData& d, uint32_t pix, float4* rec)
{
    const float4 vp = d.vtx_pos[pix], vs = d.vtx_sn[pix], vg = d.vtx_gn[pix], vv = d.vtx_view[pix];
    const pt::Material& m = d.mats[__float_as_int(vg.w)];
    const float4 thr = d.throughput[pix];
    rec[0] = make_float4(vp.x, vp.y, vp.z, vv.x);
    rec[1] = make_float4(vv.y, vv.z, vs.x, vs.y);
    rec[2] = make_float4(vs.z, m.roughness, m.base_color.r, m.base_color.g);
    rec[3] = make_float4(m.base_color.b, thr.x, thr.y, thr.z);
}

// Dense, contiguous query and training records from the two lists (right
// after k_cache_query, while the path state is that of the query vertex);
// also stores the record counts the cache kernels read.
__global__ void k_cache_compact(PtData d, CacheView cv)
{
    SegMap mq, mt;
    mq.init(cv.query_list.counts, cv.query_list.seg_cap);
    mt.init(cv.train_list.counts, cv.train_list.seg_cap);
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t tq = mq.total, tt = min(mt.total, cv.train_capacity);
    if (w == 0)
    {
        cv.counters[0] = tq;
        cv.counters[1] = tt;
    }
    const size_t sq = tq ? mq.slot(w < tq ? w : tq - 1u) : 0;
    const size_t st = tt ? mt.slot(w < tt ? w : tt - 1u) : 0;
    if (w < tq)
    {
        const uint32_t pix = cv.query_list.slots[sq];
        write_cache_record(d, pix, cv.queries + 4 * (size_t)w);
        cv.query_pix[w] = pix;
    }
    if (w < tt)
    {
        const uint32_t pix = cv.train_list.slots[st];
        write_cache_record(d, pix, cv.train_queries + 4 * (size_t)w);
        cv.train_pix[w] = pix;
        cv.train_rad[w] = d.radiance[pix];
    }
}

// radiance += throughput * max(prediction, 0) for every query.
__global__ void k_cache_resolve(PtData d, CacheView cv)
{
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= cv.counters[0])
        return;
    const uint32_t pix = cv.query_pix[row];
    const float4 thr = cv.queries[4 * (size_t)row + 3];
    const float4 o = cv.query_out[row];
    float4 r = d.radiance[pix];
    r.x += thr.y * fmaxf(o.x, 0.0f);
    r.y += thr.z * fmaxf(o.y, 0.0f);
    r.z += thr.w * fmaxf(o.z, 0.0f);
    d.radiance[pix] = r;
}

// Target of a training record: (radiance now - radiance before the query
// vertex) / throughput at it. A non-finite or negative target (from a NaN
// sample, which the framebuffer rejects) trains towards 0 and is counted.
__global__ void k_cache_targets(PtData d, CacheView cv)
{
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= cv.counters[1])
        return;
    const uint32_t pix = cv.train_pix[row];
    const float4 thr = cv.train_queries[4 * (size_t)row + 3];
    const float4 before = cv.train_rad[row], now = d.radiance[pix];
    float t[3] = {(now.x - before.x) / thr.y, (now.y - before.y) / thr.z, (now.z - before.z) / thr.w};
    bool bad = false;
    for (int k = 0; k < 3; k++)
        bad |= !(t[k] >= 0.0f) || isinf(t[k]);
    if (bad)
    {
        atomicAdd(&cv.counters[2], 1u);
        t[0] = t[1] = t[2] = 0.0f;
    }
    cv.train_targets[row] = make_float4(t[0], t[1], t[2], 0.0f);
}

inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

CacheView make_cache_view(Ctx& ctx, uint32_t sample_number, uint32_t every)
{
    Nrc& C = ctx.nrc;
    const uint32_t n = ctx.render.pixels;
    const uint32_t cap = n / every + n / (every * 8) + 1024; // expected n / every, binomial slack
    C.queries.reserve(64 * (size_t)n);
    C.query_pix.reserve(4 * (size_t)n);
    C.queries_sorted.reserve(64 * (size_t)n);
    C.query_pix_sorted.reserve(4 * (size_t)n);
    C.query_out.reserve(16 * (size_t)n);
    C.train_queries.reserve(64 * (size_t)cap);
    C.train_pix.reserve(4 * (size_t)cap);
    C.train_rad.reserve(16 * (size_t)cap);
    C.train_targets.reserve(16 * (size_t)cap);
    C.counters.reserve(16);
    const uint32_t seg_cap = seg_capacity(n);
    C.query_list.reserve(4 * (size_t)SEG_COUNT * seg_cap);
    C.train_list.reserve(4 * (size_t)SEG_COUNT * seg_cap);
    C.list_counts.reserve(4 * 2 * (size_t)SEG_COUNTER_WORDS);
    C.train_capacity = cap;
    CacheView cv{};
    cv.queries = C.queries.as<float4>();
    cv.query_pix = C.query_pix.as<uint32_t>();
    cv.query_out = C.query_out.as<float4>();
    cv.train_queries = C.train_queries.as<float4>();
    cv.train_pix = C.train_pix.as<uint32_t>();
    cv.train_rad = C.train_rad.as<float4>();
    cv.train_targets = C.train_targets.as<float4>();
    cv.counters = C.counters.as<uint32_t>();
    cv.query_list = SegList{C.query_list.as<uint32_t>(), C.list_counts.as<uint32_t>(), seg_cap};
    cv.train_list = SegList{C.train_list.as<uint32_t>(), C.list_counts.as<uint32_t>() + SEG_COUNTER_WORDS, seg_cap};
    cv.train_every = every;
    cv.train_capacity = cap;
    cv.sample_number = sample_number;
    return cv;
}

} // namespace

// MicrofacetRegularization::consistent_tau (MicrofacetRegularization.h:148-161),
// evaluated once per pass on the host in binary32.
float consistent_tau(float tau_0, uint32_t sample_number)
{
    const float pi = 3.14159265358979323846f;
    const float four_pi = 12.5663706143591729539f;
    const float two_pi = 6.28318530717958647693f;
    float p = powf((float)(int)(sample_number + 1), -1.0f / 6.0f);
    float inner = p * sqrtf(four_pi * tau_0 - 1.0f) / (two_pi * tau_0 - 1.0f);
    return 1.0f / (2.0f * pi * (1.0f - cosf(atanf(inner))));
}

PtData make_pt_data(Ctx& ctx, uint32_t sample_number, uint32_t random_number)
{
    RenderState& R = ctx.render;
    PtData d{};
    d.verts = ctx.vertices.as<float>();
    d.idx = ctx.indices.as<int32_t>();
    d.mat_idx = ctx.material_indices.as<int32_t>();
    d.mats = R.materials.as<pt::Material>();
    d.mat_class = R.mat_class.as<int32_t>();
    d.views = R.views.empty() ? nullptr : R.view_table.as<ViewDev>();
    d.view_count = (int32_t)R.views.size();
    d.frame = nullptr;
    const bool smooth = R.settings.do_normal_mapping && ctx.has_normals.ptr && ctx.vertex_normals.ptr;
    d.has_normals = smooth ? ctx.has_normals.as<uint8_t>() : nullptr;
    d.vnormals = smooth ? ctx.vertex_normals.as<float>() : nullptr;
    d.emissive = ctx.emissive_prims.as<int32_t>();
    d.areas = R.areas.as<float>();
    d.emissive_count = ctx.emissive_count;
    d.light_base = R.settings.direct_light_sampling_base_strategy;
    d.alias_probas = R.alias_probas.as<float>();
    d.alias_alias = R.alias_alias.as<int32_t>();
    d.alias_sum = R.alias_sum;
    for (int k = 0; k < 3; k++) d.cam_pos[k] = R.camera.position[k];
    d.width = R.settings.width;
    d.height = R.settings.height;
    d.nb_bounces = R.settings.nb_bounces;
    d.do_rr = R.settings.do_russian_roulette;
    d.rr_min_depth = R.settings.russian_roulette_min_depth;
    d.rr_clamp = R.settings.russian_roulette_throughput_clamp;
    d.freeze = R.settings.freeze_random;
    d.enable_direct = R.settings.enable_direct;
    d.scene_volumes = R.scene_volumes;
    d.scene_dispersive = R.scene_dispersive;
    d.sample_number = sample_number;
    d.random_number = random_number;
    d.g.ggx_conductor_ess = R.lut_conductor.as<float>();
    d.g.glossy_dielectric_ess = R.lut_glossy.as<float>();
    d.g.sheen_ltc = R.lut_sheen.as<float>();
    d.g.glass_ess = R.lut_glass.as<float>();
    d.g.glass_inv_ess = R.lut_glass_inv.as<float>();
    d.g.thin_glass_ess = R.lut_thin_glass.as<float>();
    d.g.glass_exponent = R.lut_glass_exponent.as<float>();
    d.g.mats = d.mats;
    d.g.lut_size = 128;
    d.g.consistent_tau = consistent_tau(R.settings.microfacet_regularization_tau_0, sample_number);
    d.g.consistent_tau_refract = consistent_tau(R.settings.microfacet_regularization_tau_0, sample_number + 1);
    d.g.min_roughness = R.settings.microfacet_regularization_min_roughness;
    d.g.regularize = R.settings.do_microfacet_regularization;
    d.gb_prim = R.gb_prim.as<int32_t>();
    d.gb_pos = R.gb_pos.as<float4>();
    d.gb_gn = R.gb_gn.as<uint32_t>();
    d.gb_sn = R.gb_sn.as<uint32_t>();
    d.gb_mat = R.gb_mat.as<int32_t>();
    d.throughput = R.throughput.as<float4>();
    d.radiance = R.radiance.as<float4>();
    d.vtx_pos = R.vtx_pos.as<float4>();
    d.vtx_sn = R.vtx_sn.as<float4>();
    d.vtx_gn = R.vtx_gn.as<float4>();
    d.vtx_view = R.vtx_view.as<float4>();
    d.nee_light = R.nee_light.as<float4>();
    d.nee_bsdf = R.nee_bsdf.as<float4>();
    d.rng = R.rng.as<uint32_t>();
    d.flags = R.flags.as<uint32_t>();
    d.vol_a = R.vol_a.as<uint4>();
    d.vol_b = R.vol_b.as<uint4>();
    d.shadow_rays = R.shadow_rays.as<float4>();
    d.shadow_occ = R.shadow_occ.as<uint8_t>();
    d.bsdf_rays = R.bsdf_rays.as<float4>();
    d.bsdf_hits = R.bsdf_hits.as<int4>();
    d.cont_rays = R.cont_rays.as<float4>();
    d.cont_hits = R.cont_hits.as<int4>();
    d.env = make_env_dev(ctx);
    d.env_light = R.env_light.as<float4>();
    d.env_bsdf = R.env_bsdf.as<float4>();
    d.restir = R.settings.direct_light_sampling_strategy == MIPT_LSS_RESTIR_DI;
    d.ris_all = R.settings.direct_light_sampling_strategy == MIPT_LSS_RIS_BSDF_AND_LIGHT;
    d.ris_later = d.restir && R.restir.later_bounces_sampling_strategy == MIPT_RESTIR_DI_LATER_BOUNCES_RIS_BSDF_AND_LIGHT;
    d.ris_light = R.settings.ris_number_of_light_candidates;
    d.ris_bsdf = R.settings.ris_number_of_bsdf_candidates;
    d.ris_state = R.ris_state.as<float4>();
    d.final_shading_visibility = R.restir.do_final_shading_visibility;
    d.restir_out = d.restir ? restir_reservoir_buffer(ctx, R.restir_shade_buffer) : nullptr;
    d.fb = R.fb.as<float>();
    d.fb64 = R.settings.accumulate_fp64 ? R.fb64.as<double>() : nullptr;
    d.rejected = R.counters.as<uint32_t>();
    return d;
}

// The per-sample record of a graph replay, written by one thread: a
// kernel's arguments are copied at launch, so the host may fill the next
// sample's values at once (no pinned buffer to keep alive).
__global__ void k_set_frame(FrameParams* __restrict__ frame, FrameParams values) { *frame = values; }

void write_frame_params(FrameParams* frame, const FrameParams& values, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_frame, dim3(1), dim3(1), 0, s, frame, values);
    HIP_CHECK(hipGetLastError());
}

void camera_rays_pass(Ctx& ctx, uint32_t sample_number, uint32_t random_number, hipStream_t s,
                      const FrameParams* frame)
{
    RenderState& R = ctx.render;
    const int32_t W = R.settings.width, H = R.settings.height;
    const uint32_t n = (uint32_t)W * (uint32_t)H;
    const bool keep_prev = restir_needs_prev_gbuffer(R);
    GBufferView cur{R.gb_prim.as<int32_t>(), R.gb_pos.as<float4>(), R.gb_gn.as<uint32_t>(), R.gb_sn.as<uint32_t>(),
                    R.gb_mat.as<int32_t>()};
    GBufferView prev{R.pg_prim.as<int32_t>(), R.pg_pos.as<float4>(), R.pg_gn.as<uint32_t>(), R.pg_sn.as<uint32_t>(),
                     R.pg_mat.as<int32_t>()};
    // prev_camera: the camera of the previous frame (the first frame's own)
    R.prev_camera = R.frame_camera_valid ? R.frame_camera : R.camera;
    R.frame_camera = R.camera;
    R.frame_camera_valid = true;
    if (sample_number == 0 && !frame && R.settings.accumulate)
    {
        // need_to_reset: G-buffers and ReSTIR DI reservoirs start over; not
        // accumulating, a moving camera does not reset them (GPURenderer::reset).
        hipLaunchKernelGGL(k_gbuffer_reset, dim3(blocks_for(n, 256)), dim3(256), 0, s, cur, n);
        if (keep_prev)
            hipLaunchKernelGGL(k_gbuffer_reset, dim3(blocks_for(n, 256)), dim3(256), 0, s, prev, n);
        restir_reset(ctx, s);
    }
    if (!R.views.empty())
        camera_rays_views(ctx, sample_number, random_number, R.settings.freeze_random, R.cont_rays.as<mipt_ray>(), s,
                          frame);
    else
        camera_rays(ctx, R.camera, W, H, sample_number, random_number, R.settings.freeze_random,
                    R.cont_rays.as<mipt_ray>(), nullptr, s, frame);
    trace_closest(ctx, R.cont_rays.as<mipt_ray>(), R.cont_hits.as<mipt_hit>(), nullptr, n, s);
    PtData d = make_pt_data(ctx, sample_number, random_number);
    d.frame = frame;
    hipLaunchKernelGGL(k_gbuffer_write, dim3(blocks_for(n, 256)), dim3(256), 0, s, d, d.cont_rays, d.cont_hits, prev,
                       (int32_t)keep_prev);
    HIP_CHECK(hipGetLastError());
}

void path_trace_pass(Ctx& ctx, uint32_t sample_number, uint32_t random_number, hipStream_t s, bool warmup,
                     const FrameParams* frame)
{
    RenderState& R = ctx.render;
    const uint32_t n = (uint32_t)R.settings.width * (uint32_t)R.settings.height;
    PtData d = make_pt_data(ctx, sample_number, random_number);
    d.frame = frame;
    const unsigned g = blocks_for(n, 256);
    const bool cache = R.settings.radiance_cache != 0;
    if (warmup && !cache)
        fail(MIPT_ERR_INVALID_ARGUMENT, "RADIANCE_CACHE_WARMUP needs render_settings.radiance_cache = 1");
    CacheView cv{};
    if (cache)
    {
        if (!ctx.nrc.configured)
            fail(MIPT_ERR_INVALID_ARGUMENT, "radiance_cache is on: call mipt_set_radiance_cache first");
        // A warm-up pass traces only its training paths, one in
        // warmup_train_every, and leaves the framebuffer alone.
        const uint32_t every =
            (uint32_t)(warmup ? ctx.nrc.settings.warmup_train_every : ctx.nrc.settings.train_every);
        d.train_only_every = warmup ? every : 0u;
        cv = make_cache_view(ctx, sample_number, every);
        hipLaunchKernelGGL(k_cache_counters_reset, dim3(1), dim3(64), 0, s, cv.counters);
    }
    const uint32_t seg_cap = seg_capacity(n);
    uint32_t* qc = R.q_counts.as<uint32_t>();
    const size_t qn = (size_t)SEG_COUNT * seg_cap;
    uint32_t* qskip = R.q_skip.as<uint32_t>();
    // shadow rays: the light sample's, and with envmap sampling two more per
    // path (k_shade appends up to three times per thread)
    const uint32_t shadow_slots = env_sampling_on(ctx) ? 3u : 1u;
    if (R.shadow_slots < shadow_slots)
        fail(MIPT_ERR_INVALID_ARGUMENT, "envmap sampling: the shadow ray buffers were not sized for it");
    const RayQueues q{SegList{R.q_shadow.as<uint32_t>(), qc, shadow_slots * seg_cap},
                      SegList{R.q_bsdf.as<uint32_t>(), qc + SEG_COUNTER_WORDS, seg_cap},
                      SegList{R.q_cont.as<uint32_t>(), qc + 2 * SEG_COUNTER_WORDS, seg_cap},
                      {SegList{qskip, qc + 3 * SEG_COUNTER_WORDS, seg_cap},
                       SegList{qskip + qn, qc + 4 * SEG_COUNTER_WORDS, seg_cap}}};
    // Live paths of a bounce, ping-ponged: the list a bounce runs over and
    // the one its resolve kernels fill for the next. k_resolve and each skip
    // round append to it, each at most once per thread, so a segment holds up
    // to LIVE_APPENDERS times a one-kernel list's share.
    uint32_t* qlive = R.q_live.as<uint32_t>();
    const uint32_t live_cap = LIVE_APPENDERS * seg_cap;
    const SegList live[2] = {SegList{qlive, qc + 5 * SEG_COUNTER_WORDS, live_cap},
                             SegList{qlive + (size_t)SEG_COUNT * live_cap, qc + 6 * SEG_COUNTER_WORDS, live_cap}};
    // A skip needs a medium of higher priority than the boundary hit; only
    // transmissive materials take priorities below the maximum.
    const int32_t skip_rounds = d.scene_volumes ? MAX_SKIP_ROUNDS : 0;
    float4* thr_prev = R.thr_prev.as<float4>();
    // Sorting pays where paths inside media and at opaque surfaces share
    // waves (a nested-glass scene: 13.26 -> 12.59 ms per 1080p sample); on
    // opaque scenes it costs its two kernels (config 1: 9.03 -> 9.32 ms).
    const bool sort_paths = R.shade_order == MIPT_SHADE_ORDER_MATERIAL ||
                            (R.shade_order == MIPT_SHADE_ORDER_AUTO && R.scene_volumes);
    hipLaunchKernelGGL(k_seg_reset, dim3(1), dim3(SEG_COUNT), 0, s, live[0].counts, 1u);
    hipLaunchKernelGGL(k_path_init, dim3(g), dim3(256), 0, s, d, live[0]);
    int cur = 0;
    SegList cur_list = live[0];
    for (int32_t bounce = 0; bounce <= R.settings.nb_bounces; bounce++)
    {
        if (cache && bounce == ctx.nrc.settings.query_bounce)
        {
            hipLaunchKernelGGL(k_seg_reset, dim3(1), dim3(2 * SEG_COUNT), 0, s, cv.query_list.counts, 2u);
            hipLaunchKernelGGL(k_cache_query, dim3(g), dim3(256), 0, s, d, cv, cur_list);
            hipLaunchKernelGGL(k_cache_compact, dim3(g), dim3(256), 0, s, d, cv);
            cur_list = cv.train_list; // the paths that go on
        }
        const SegList next = live[cur ^ 1];
        hipLaunchKernelGGL(k_seg_reset, dim3(1), dim3(4 * SEG_COUNT), 0, s, qc, 4u);
        hipLaunchKernelGGL(k_seg_reset, dim3(1), dim3(SEG_COUNT), 0, s, next.counts, 1u);
        const uint32_t* sorted = nullptr;
        if (sort_paths)
        {
            uint32_t* sc = R.sort_counts.as<uint32_t>();
            uint32_t* cursors = sc + (size_t)SHADE_SORT_COUNTERS * SEG_STRIDE;
            hipLaunchKernelGGL(k_sort_reset, dim3(blocks_for(2 * SHADE_SORT_COUNTERS, 256)), dim3(256), 0, s, sc);
            hipLaunchKernelGGL(k_sort_hist, dim3(g), dim3(256), 0, s, d, cur_list, sc);
            hipLaunchKernelGGL(k_sort_scatter, dim3(g), dim3(256), 0, s, d, cur_list, (const uint32_t*)sc, cursors,
                               R.shade_sorted.as<uint32_t>());
            sorted = R.shade_sorted.as<uint32_t>();
        }
        hipLaunchKernelGGL(k_shade, dim3(g), dim3(256), 0, s, d, bounce, q, cur_list, sorted, thr_prev);
        // Shadow (any hit), BSDF-to-light and continuation rays. At bounce 0
        // every pixel has its rays: one launch per kind, at the single-kind
        // kernel's higher occupancy. Later bounces carry fewer paths, where a
        // launch costs about its longest ray's latency: one launch for all.
        const bool cont_rays = bounce < R.settings.nb_bounces;
        if (bounce == 0)
        {
            trace_any_queue(ctx, R.shadow_rays.as<mipt_ray>(), R.shadow_occ.as<uint8_t>(), q.shadow,
                            shadow_slots * n, s);
            trace_closest_queue(ctx, R.bsdf_rays.as<mipt_ray>(), R.bsdf_hits.as<mipt_hit>(), q.bsdf, n, s);
            if (cont_rays)
                trace_closest_queue(ctx, R.cont_rays.as<mipt_ray>(), R.cont_hits.as<mipt_hit>(), q.cont, n, s);
        }
        else
        {
            const TraceQueue tq[3] = {
                {R.shadow_rays.as<mipt_ray>(), nullptr, R.shadow_occ.as<uint8_t>(), &q.shadow},
                {R.bsdf_rays.as<mipt_ray>(), R.bsdf_hits.as<mipt_hit>(), nullptr, &q.bsdf},
                {R.cont_rays.as<mipt_ray>(), R.cont_hits.as<mipt_hit>(), nullptr, &q.cont}};
            trace_queues(ctx, tq, cont_rays ? 3 : 2, n, s);
        }
        hipLaunchKernelGGL(k_resolve, dim3(g), dim3(256), 0, s, d, bounce, (const float4*)thr_prev, cur_list, q.skip[0],
                           next);
        for (int32_t k = 0; k < skip_rounds && bounce < R.settings.nb_bounces; k++)
        {
            const SegList& in = q.skip[k & 1];
            const SegList& out = q.skip[(k + 1) & 1];
            hipLaunchKernelGGL(k_seg_reset, dim3(1), dim3(SEG_COUNT), 0, s, out.counts, 1u);
            trace_closest_queue(ctx, R.cont_rays.as<mipt_ray>(), R.cont_hits.as<mipt_hit>(), in, n, s);
            hipLaunchKernelGGL(k_resolve_skip, dim3(g), dim3(256), 0, s, d, in, out, next);
        }
        cur ^= 1;
        cur_list = next;
    }
    if (cache && !warmup)
    {
        // Queries in spatial order: a wave's rows then share hash-grid
        // entries, and the encode's gathers hit the caches.
        CacheView sorted = cv;
        sorted.queries = ctx.nrc.queries_sorted.as<float4>();
        sorted.query_pix = ctx.nrc.query_pix_sorted.as<uint32_t>();
        nrc_spatial_order(ctx, cv.queries, cv.query_pix, cv.counters, n, sorted.queries, sorted.query_pix, s);
        nrc_infer(ctx, reinterpret_cast<const float*>(sorted.queries), cv.counters, n,
                  reinterpret_cast<float*>(ctx.nrc.query_out.ptr), s);
        hipLaunchKernelGGL(k_cache_resolve, dim3(g), dim3(256), 0, s, d, sorted);
    }
    if (cache)
        hipLaunchKernelGGL(k_cache_targets, dim3(blocks_for(cv.train_capacity, 256)), dim3(256), 0, s, d, cv);
    if (!warmup)
        hipLaunchKernelGGL(k_accumulate, dim3(g), dim3(256), 0, s, d);
    HIP_CHECK(hipGetLastError());
}

} // namespace mipt
