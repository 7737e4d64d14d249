// restir.hip -- ReSTIR DI passes on the wavefront design.
//
// Each reference kernel is one per-pixel function that traces rays inline
// (src/Device/kernels/ReSTIR/DI/InitialCandidates.h:478-538,
// TemporalReuse.h:50-296, SpatialReuse.h:43-322). Here every ray query is a
// batched traversal launch between two kernels, and the per-pixel state
// (reservoir in progress, xorshift32 state) lives in a scratch buffer in
// between, so the RNG draws happen in the reference's order:
//
//   INITIAL  k_ic_start (light candidates, first BSDF candidate ray)
//            -> closest hit -> k_ic_bsdf (x number_of_initial_bsdf_candidates)
//            -> any hit (visibility reuse) -> k_ic_finish
//   TEMPORAL k_tp_start (target functions without visibility, 2 shadow rays)
//            -> any hit -> k_tp_finish (pairwise MIS, two combines)
//   SPATIAL  k_sp_start (neighbour choice and heuristics, neighbour combines,
//            one shadow ray per neighbour for the canonical MIS weight)
//            -> any hit -> k_sp_center (canonical combine, normalisation)
//            -> any hit (visibility reuse) -> k_sp_finish
//
// Fixed kernel options (ReSTIRDIOptions.h defaults): pairwise MIS with
// visibility in the bias correction, visibility reuse, no visibility in the
// initial or spatial target functions, no light presampling, no envmap, the
// camera static between frames (accumulation: the temporal neighbour is the
// pixel itself, UtilsTemporal.h:37-41).
//
// Reservoir record (ReSTIRDIReservoir without weight_sum, which no later pass
// reads): float4 (point_on_light_source, target_function),
// float4 (M bits, emissive_triangle_index bits, UCW, flags bits).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "pt_device.h"

namespace mipt {
namespace {

constexpr int32_t PHASE_SKIP = 0;     // no primary hit: the reference returns without writing
constexpr int32_t PHASE_PENDING = 1;  // a BSDF candidate ray is in flight
constexpr int32_t PHASE_DONE = 2;     // reservoir finished, visibility ray (maybe) in flight
constexpr int32_t PHASE_COPY = 3;     // temporal: output = initial candidates

struct Res {
    int32_t M;
    float wsum;
    float ucw;
    f3 point;
    int32_t tri;
    float target;
    uint32_t flags;
};

__device__ __forceinline__ Res res_empty()
{
    Res r;
    r.M = 0;
    r.wsum = 0.0f;
    r.ucw = 0.0f;
    r.point = pt::mk(0.0f, 0.0f, 0.0f);
    r.tri = -1;
    r.target = 0.0f;
    r.flags = 0u;
    return r;
}

__device__ __forceinline__ Res res_load(const float4* buf, size_t pix)
{
    const float4 a = buf[2 * pix], b = buf[2 * pix + 1];
    Res r;
    r.point = ld3(a);
    r.target = a.w;
    r.M = __float_as_int(b.x);
    r.tri = __float_as_int(b.y);
    r.ucw = b.z;
    r.flags = (uint32_t)__float_as_int(b.w);
    r.wsum = 0.0f;
    return r;
}

__device__ __forceinline__ void res_store(float4* buf, size_t pix, const Res& r)
{
    buf[2 * pix] = st3(r.point, r.target);
    buf[2 * pix + 1] = make_float4(__int_as_float(r.M), __int_as_float(r.tri), r.ucw, __int_as_float((int32_t)r.flags));
}

// Scratch state of a pixel between kernels: 4 float4.
struct RsData {
    float4* presampled; // presampled lights: 3 float4 each (point, tri), (normal, pdf), (radiance, flags)
    float4* state;      // 4 per pixel
    float4* rays;       // slots of 2 float4, slot k of pixel p at k * n + p
    const uint8_t* occ; // per slot
    const int4* hits;   // per pixel
    float* aux;         // 2 * K per pixel (spatial)
    const float4* in;   // reservoirs read
    float4* out;        // reservoirs written
    const float4* init; // initial candidates (temporal)
    // previous-frame G-buffer (temporal)
    const int32_t* pg_prim;
    const float4* pg_pos;
    const uint32_t* pg_gn;
    const uint32_t* pg_sn;
    const int32_t* pg_mat;
    uint32_t n;
    mipt_restir_di_settings s;
    int32_t pass_index;
    int32_t kill_after; // spatial: visibility reuse after this pass
    // the previous frame's camera (temporal reprojection when not accumulating)
    int32_t accumulate;
    float prev_vp[4][4];
    float prev_cam_pos[3];
};

__device__ __forceinline__ void state_store(const RsData& r, uint32_t pix, const Res& res, uint32_t rng, int32_t phase,
                                            float x0, float x1, uint32_t x2)
{
    float4* s = r.state + 4 * (size_t)pix;
    s[0] = st3(res.point, res.target);
    s[1] = make_float4(__int_as_float(res.M), __int_as_float(res.tri), res.wsum, __int_as_float((int32_t)res.flags));
    s[2] = make_float4(res.ucw, __int_as_float((int32_t)rng), __int_as_float(phase), x0);
    s[3] = make_float4(x1, __int_as_float((int32_t)x2), 0.0f, 0.0f);
}

__device__ __forceinline__ void state_load(const RsData& r, uint32_t pix, Res& res, uint32_t& rng, int32_t& phase,
                                           float& x0, float& x1, uint32_t& x2)
{
    const float4* s = r.state + 4 * (size_t)pix;
    const float4 a = s[0], b = s[1], c = s[2], d = s[3];
    res.point = ld3(a);
    res.target = a.w;
    res.M = __float_as_int(b.x);
    res.tri = __float_as_int(b.y);
    res.wsum = b.z;
    res.flags = (uint32_t)__float_as_int(b.w);
    res.ucw = c.x;
    rng = (uint32_t)__float_as_int(c.y);
    phase = __float_as_int(c.z);
    x0 = c.w;
    x1 = d.x;
    x2 = (uint32_t)__float_as_int(d.y);
}

// ReSTIRDIReservoir::add_one_candidate (DI/Reservoir.h:54-61)
__device__ __forceinline__ void add_candidate(Res& r, f3 point, int32_t tri, float target, uint32_t flags, float weight,
                                              pt::Xorshift32& rng)
{
    r.M++;
    r.wsum += weight;
    if (rng() < weight / r.wsum)
    {
        r.point = point;
        r.tri = tri;
        r.target = target;
        r.flags = flags;
    }
}

// combine_with (DI/Reservoir.h:77-101)
__device__ __forceinline__ bool combine_with(Res& r, const Res& other, float mis, float target, float jacobian,
                                             pt::Xorshift32& rng)
{
    if (other.ucw <= 0.0f)
    {
        r.M += other.M;
        return false;
    }
    const float w = mis * target * other.ucw * jacobian;
    r.M += other.M;
    r.wsum += w;
    if (rng() < w / r.wsum)
    {
        r.point = other.point;
        r.tri = other.tri;
        r.flags = other.flags;
        r.target = target;
        return true;
    }
    return false;
}

// end (DI/Reservoir.h:103-109)
__device__ __forceinline__ void res_end(Res& r) { r.ucw = r.wsum == 0.0f ? 0.0f : 1.0f / r.target * r.wsum; }

// end_with_normalization (DI/Reservoir.h:111-121)
__device__ __forceinline__ void res_end_normalized(Res& r, float num, float den)
{
    if (r.wsum == 0.0f || r.wsum < 1.0e-10f || r.wsum > 1.0e10f || den == 0.0f || num == 0.0f)
        r.ucw = 0.0f;
    else
        r.ucw = 1.0f / r.target * r.wsum * num / den;
    r.M = r.M < 1000000 ? r.M : 1000000;
}

// ReSTIRSurface (ReSTIR/Surface.h:13-41). The view direction is taken from
// the current frame's hit position also for a previous-frame surface
// (Surface.h:62, get_view_direction of the current G-buffer).
struct Surface {
    int32_t mat;
    int32_t prim;
    f3 view, sn, gn, p;
};

// The surface's volume state (RayVolumeState::reconstruct_first_hit,
// RayVolumeState.h:35-59): its material pushed on an empty stack. A surface
// without a primary hit keeps the empty stack (the early return at
// RayVolumeState.h:37-39): a previous-frame pixel whose camera ray missed has
// prim -1 and the material of an older hit, and reaches the temporal target
// function. Its incident and outgoing media are the air: the reference leaves
// them at -1 (RayVolumeState.h:65) and its BSDF evaluation then reads the IOR
// of material -1 (Principled.h:1646), an out-of-bounds read. Dispersive
// glass, which would also draw a wavelength, is refused with ReSTIR DI.
__device__ __forceinline__ pt::VolumeState surface_volume(const PtData& d, const Surface& s)
{
    pt::VolumeState v;
    pt::vol_init(v);
    if (s.prim >= 0)
        pt::vol_push(v, d.mat_idx[s.prim], d.mats[s.mat].dielectric_priority);
    else
        v.incident = v.outgoing = (int32_t)pt::VOL_AIR;
    return v;
}

__device__ __forceinline__ Surface surface_current(const PtData& d, uint32_t pix)
{
    Surface s;
    s.mat = d.gb_mat[pix];
    s.prim = d.gb_prim[pix];
    const f3 cam = pt::mk(d.cam_pos[0], d.cam_pos[1], d.cam_pos[2]);
    s.p = ld3(d.gb_pos[pix]);
    s.view = pt::normalize(cam - s.p);
    s.sn = gb_normal(d.gb_sn[pix]);
    s.gn = gb_normal(d.gb_gn[pix]);
    return s;
}

// get_pixel_surface_previous_frame (Surface.h:55-71): the previous G-buffer,
// seen from the previous frame's camera through the current frame's hit
// position at that pixel.
__device__ __forceinline__ Surface surface_previous(const PtData& d, const RsData& r, uint32_t pix)
{
    Surface s;
    s.mat = r.pg_mat[pix];
    s.prim = r.pg_prim[pix];
    const f3 cam = pt::mk(r.prev_cam_pos[0], r.prev_cam_pos[1], r.prev_cam_pos[2]);
    s.view = pt::normalize(cam - ld3(d.gb_pos[pix]));
    s.sn = gb_normal(r.pg_sn[pix]);
    s.gn = gb_normal(r.pg_gn[pix]);
    s.p = ld3(r.pg_pos[pix]);
    return s;
}

// Direction and distance from p to a sample: an envmap sample stores its
// direction in envmap space and lies at 1e35 (ReSTIR_DI_get_light_sample_
// direction, DI/TargetFunction.h:13-29).
__device__ __forceinline__ f3 sample_direction(const PtData& d, f3 p, f3 point, uint32_t flags, float& dist)
{
    if (flags & pt::RESTIR_FLAG_ENVMAP)
    {
        dist = 1.0e35f;
        return env_rotate(d.env.to_world, point);
    }
    const f3 to = point - p;
    dist = sqrtf(pt::dot(to, to));
    return pt::vdiv(to, dist);
}

// ReSTIR_DI_evaluate_target_function<false> (DI/TargetFunction.h:49-98): the
// luminance of f * Le * cos * G (G = 1 for an envmap sample, whose Le is the
// envmap's radiance in that direction). The visibility factor, when wanted,
// is applied by the caller from a shadow ray (origin s.p, direction dir_out,
// tmax dist_out - 1e-4), traced only when the value is non-zero.
__device__ __forceinline__ float target_function(const PtData& d, const Surface& s, f3 point, int32_t tri, uint32_t flags,
                                                 f3& dir_out, float& dist_out)
{
    dist_out = 0.0f;
    const bool env = (flags & pt::RESTIR_FLAG_ENVMAP) != 0u;
    if ((tri == -1 && !env) || s.mat < 0)
        return 0.0f;
    float dist;
    const f3 dir = sample_direction(d, s.p, point, flags, dist);
    dir_out = dir;
    dist_out = dist;
    const float cosine = pt::dot(s.sn, dir);
    if (cosine <= 0.0f)
        return 0.0f;
    const pt::Material& m = d.mats[s.mat];
    pt::VolumeState vs = surface_volume(d, s);
    pt::BsdfCtx c{s.view, s.sn, s.gn, dir, flags & pt::RESTIR_FLAG_LOBE_MASK, pt::NO_REGULARIZATION, 0.0f, &vs, 0};
    float bpdf;
    const Color f = pt::principled_eval(d.g, m, c, bpdf);
    Color emission;
    float geometry = 1.0f;
    if (env)
    {
        float epdf;
        emission = env_eval(d.env, dir, epdf);
    }
    else
    {
        emission = d.mats[d.mat_idx[tri]].emission;
        geometry = cos_at_light(pt::normalize(tri_normal_unnormalized(d, tri)), -1.0f * dir);
        geometry /= dist * dist;
    }
    return luminance(((f * emission) * cosine) * geometry);
}

// The probability that a light candidate samples the envmap rather than an
// emissive triangle (InitialCandidates.h:451-459, LightsPresampling.h:99-107);
// an envmap of intensity 0 counts as no envmap.
__device__ __forceinline__ float env_candidate_probability(const PtData& d, const RsData& r)
{
    if (!env_sampled(d.env))
        return 0.0f;
    return d.emissive_count == 0 ? 1.0f : r.s.envmap_candidate_probability;
}

// target with visibility from a traced shadow ray; NaN stays NaN as in
// `target *= visible`.
__device__ __forceinline__ float apply_visibility(float target, bool occluded) { return target * (occluded ? 0.0f : 1.0f); }

__device__ __forceinline__ bool is_emissive(const PtData& d, int32_t mat)
{
    if (mat < 0)
        return false;
    const Color e = d.mats[mat].emission;
    const float nz = 1.0e-10f;
    return !(e.r < nz && e.r > -nz) || !(e.g < nz && e.g > -nz) || !(e.b < nz && e.b > -nz);
}

// ---------------------------------------------------------------------------
// LIGHT PRESAMPLING (LightsPresampling.h:74-112) and its use by the initial
// candidates
// ---------------------------------------------------------------------------
// cantor_pairing_function (InitialCandidates.h:28-31)
__device__ __forceinline__ int32_t cantor_pair(int32_t x, int32_t y) { return (x + y + 1) * (x + y) / 2 + y; }

// Index of the presampled light that light candidate draws for pixel `pix`:
// subset from the tile's stream (seeded random_number * (tile + 1)), light of
// the subset from the pixel's stream (one draw).
__device__ __forceinline__ uint32_t presampled_index(const PtData& d, const RsData& r, uint32_t pix, pt::Xorshift32& rng)
{
    const int32_t ts = r.s.light_presampling_tile_size;
    const int32_t x = (int32_t)(pix % (uint32_t)d.width), y = (int32_t)(pix / (uint32_t)d.width);
    const int32_t tile = cantor_pair(x / ts, y / ts);
    pt::Xorshift32 subset_rng{d.random_number * (uint32_t)(tile + 1)};
    const int32_t subset = subset_rng.random_index(r.s.light_presampling_number_of_subsets);
    const int32_t in_subset = rng.random_index(r.s.light_presampling_subset_size);
    return (uint32_t)subset * (uint32_t)r.s.light_presampling_subset_size + (uint32_t)in_subset;
}

// One presampled light per thread: an envmap draw (probability 0 here), then
// an emissive triangle by power (ReSTIR_DI_LightPresamplingStrategy,
// ReSTIRDIOptions.h:162). A zero-pdf sample keeps the empty record.
__global__ void k_presample_lights(PtData d, float env_prob, float4* __restrict__ out, uint32_t count)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= count)
        return;
    pt::Xorshift32 rng{pt::pixel_seed(x, d.sample_number, d.random_number, d.freeze != 0)};
    float4 a = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
    float4 b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 c = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(0));
    if (rng() < env_prob)
    {
        // presample_envmap (LightsPresampling.h:22-35): the direction in
        // envmap space
        f3 dir;
        float pdf;
        const Color radiance = env_sample(d.env, dir, pdf, rng);
        a = st3(env_rotate(d.env.to_env, dir), __int_as_float(-1));
        b = st3(pt::mk(0.0f, 0.0f, 0.0f), pdf * env_prob);
        c = stc(radiance, __int_as_float((int32_t)pt::RESTIR_FLAG_ENVMAP));
    }
    else
    {
        const LightSample ls = sample_emissive_with(d, MIPT_LSS_BASE_POWER, rng);
        if (ls.area_pdf > 0.0f)
        {
            a = st3(ls.point, __int_as_float(ls.tri));
            b = st3(ls.normal, ls.area_pdf * (1.0f - env_prob));
            c = stc(ls.emission, __int_as_float(0));
        }
    }
    out[3 * (size_t)x] = a;
    out[3 * (size_t)x + 1] = b;
    out[3 * (size_t)x + 2] = c;
}

// ---------------------------------------------------------------------------
// INITIAL CANDIDATES
// ---------------------------------------------------------------------------

// Samples BSDF candidates from index `cand` on until one needs a ray
// (returns true, ray written to slot 0) or none is left (false).
__device__ __forceinline__ bool next_bsdf_candidate(const PtData& d, const RsData& r, uint32_t pix, const Surface& s,
                                                    const pt::Material& m, int32_t& cand, pt::Xorshift32& rng,
                                                    float& bpdf_out, uint32_t& info_out)
{
    const int32_t nb = r.s.number_of_initial_bsdf_candidates;
    while (cand < nb)
    {
        pt::VolumeState vs = surface_volume(d, s);
        pt::BsdfCtx c{s.view, s.sn, s.gn, pt::mk(0.0f, 0.0f, 0.0f), pt::NO_INFO, pt::NO_REGULARIZATION, 0.0f, &vs, 0};
        f3 dir;
        float bpdf;
        pt::principled_sample(d.g, m, c, dir, bpdf, rng);
        if (bpdf > 0.0f)
        {
            write_ray(r.rays, pix, s.p, s.prim, dir, 1.0e35f - 1.0e-4f);
            bpdf_out = bpdf;
            info_out = c.info;
            return true;
        }
        cand++;
    }
    return false;
}

// reservoir.end(), M = 1 (InitialCandidates.h:465-472), then the visibility
// reuse test (ReSTIR/Utils.h:25-67) as a ray in slot 1.
__device__ __forceinline__ void finish_initial(const PtData& d, const RsData& r, uint32_t pix, const Surface& s,
                                               Res& res, uint32_t& need_vis)
{
    res_end(res);
    res.M = 1;
    need_vis = 0u;
    write_inactive(r.rays, (size_t)r.n + pix);
    if (res.ucw <= 0.0f && res.wsum <= 0.0f)
        return;
    if (res.flags & pt::RESTIR_FLAG_UNOCCLUDED)
        return;
    float dist;
    const f3 dir = sample_direction(d, s.p, res.point, res.flags, dist);
    write_ray(r.rays, (size_t)r.n + pix, s.p, s.prim, dir, dist - 1.0e-4f);
    need_vis = 1u;
}

__global__ __launch_bounds__(256, 2) void k_ic_start(PtData d, RsData r)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= r.n)
        return;
    write_inactive(r.rays, pix);
    pt::Xorshift32 rng{pt::pixel_seed(pix, d.sample_number, d.random_number, d.freeze != 0)};
    if (d.gb_prim[pix] == -1)
    {
        write_inactive(r.rays, (size_t)r.n + pix);
        state_store(r, pix, res_empty(), 0u, PHASE_SKIP, 0.0f, 0.0f, 0u);
        return;
    }
    const Surface s = surface_current(d, pix);
    const pt::Material& m = d.mats[s.mat];
    Res res = res_empty();
    const int32_t nl = r.s.number_of_initial_light_candidates;
    const int32_t nb = r.s.number_of_initial_bsdf_candidates;
    const float env_prob = env_candidate_probability(d, r);
    // sample_light_candidates (InitialCandidates.h:180-298)
    for (int32_t i = 0; i < nl; i++)
    {
        float cosine = 0.0f, pdf = 0.0f, dist = 0.0f;
        f3 dir = pt::mk(0.0f, 0.0f, 0.0f);
        f3 point = pt::mk(0.0f, 0.0f, 0.0f);
        f3 ls_normal = pt::mk(0.0f, 0.0f, 0.0f);
        Color ls_emission = pt::col(0.0f);
        int32_t tri = -1;
        uint32_t sflags = 0u;
        if (r.presampled)
        {
            // use_presampled_light_candidate (InitialCandidates.h:33-90): the
            // pixel's tile picks a subset with its own stream, the pixel a
            // light of the subset
            const float4* pl = r.presampled + 3 * (size_t)presampled_index(d, r, pix, rng);
            point = ld3(pl[0]);
            tri = __float_as_int(pl[0].w);
            ls_normal = ld3(pl[1]);
            pdf = pl[1].w;
            ls_emission = ldc(pl[2]);
            sflags = (uint32_t)__float_as_int(pl[2].w);
            dir = sample_direction(d, s.p, point, sflags, dist);
            cosine = pt::dot(s.sn, dir);
        }
        // sample_fresh_light_candidate (:92-178)
        else if (rng() > env_prob)
        {
            const LightSample ls = sample_emissive(d, rng);
            tri = ls.tri;
            point = ls.point;
            pdf = ls.area_pdf;
            ls_normal = ls.normal;
            ls_emission = ls.emission;
            if (pdf > 0.0f)
            {
                const f3 to = point - s.p;
                dist = sqrtf(pt::dot(to, to));
                dir = pt::vdiv(to, dist);
                cosine = pt::maxf_(0.0f, pt::dot(s.sn, dir));
                pdf *= (1.0f - env_prob);
            }
            else
            {
                const f3 to = point - s.p;
                dist = sqrtf(pt::dot(to, to));
                dir = pt::vdiv(to, dist);
            }
        }
        // A draw of exactly 0 (the zero xorshift state that a seed of
        // wang_hash(61) = 0 gives under freeze_random) lands here with
        // probability 0 too: without an envmap it stays no candidate.
        else if (env_prob > 0.0f)
        {
            // an envmap candidate, stored as its direction in envmap space;
            // the candidate's own direction is recomputed from that
            f3 edir;
            ls_emission = env_sample(d.env, edir, pdf, rng);
            cosine = pt::maxf_(0.0f, pt::dot(edir, s.sn));
            pdf *= env_prob;
            point = env_rotate(d.env.to_env, edir);
            sflags = pt::RESTIR_FLAG_ENVMAP;
            dir = sample_direction(d, s.p, point, sflags, dist);
        }
        const bool env = (sflags & pt::RESTIR_FLAG_ENVMAP) != 0u;
        float weight = 0.0f, sample_target = 0.0f;
        if (cosine > 0.0f && pdf > 0.0f)
        {
            pt::VolumeState vs = surface_volume(d, s);
            pt::BsdfCtx c{s.view, s.sn, s.gn, dir, pt::NO_INFO, pt::NO_REGULARIZATION, 0.0f, &vs, 0};
            float bsdf_pdf;
            const Color f = pt::principled_eval(d.g, m, c, bsdf_pdf);
            // ReSTIR_DI_evaluate_target_function<false> of this light sample:
            // its BSDF evaluation has the same arguments as the one above (the
            // direction is recomputed from the same point by the same
            // operations), and the normalised triangle normal equals the light
            // sample's normal bit for bit, so both are reused. An envmap
            // sample's emission is the envmap evaluated in its direction.
            float target;
            if (env)
            {
                const float tcos = pt::dot(s.sn, dir);
                float epdf;
                const Color e = env_eval(d.env, dir, epdf);
                target = tcos <= 0.0f ? 0.0f : luminance(((f * e) * tcos) * 1.0f);
            }
            else
            {
                const float geometry = cos_at_light(ls_normal, -1.0f * dir) / (dist * dist);
                target = luminance(((f * ls_emission) * cosine) * geometry);
            }
            if (!(bsdf_pdf <= 0.0f))
            {
                const f3 ln = ls_normal;
                float lpdf = pdf;
                if (!env)
                {
                    lpdf = area_to_solid_angle(pdf / (1.0f - env_prob), dist, cos_at_light(ln, -1.0f * dir));
                    lpdf *= (1.0f - env_prob);
                }
                const float mis = balance_heuristic(lpdf, (float)nl, bsdf_pdf, (float)nb);
                weight = mis * target / pdf;
                sample_target = target;
            }
        }
        add_candidate(res, point, tri, sample_target, sflags, weight, rng);
    }
    int32_t cand = 0;
    float bpdf = 0.0f;
    uint32_t info = 0u;
    if (next_bsdf_candidate(d, r, pix, s, m, cand, rng, bpdf, info))
    {
        write_inactive(r.rays, (size_t)r.n + pix);
        state_store(r, pix, res, rng.state, PHASE_PENDING, bpdf, __int_as_float(cand), info);
        return;
    }
    uint32_t need_vis;
    finish_initial(d, r, pix, s, res, need_vis);
    state_store(r, pix, res, rng.state, PHASE_DONE, 0.0f, 0.0f, need_vis);
}

// sample_bsdf_candidates (InitialCandidates.h:300-440) after the closest hit.
__global__ __launch_bounds__(256, 2) void k_ic_bsdf(PtData d, RsData r)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= r.n)
        return;
    Res res;
    uint32_t rng_state, info;
    int32_t phase;
    float bpdf, candf;
    state_load(r, pix, res, rng_state, phase, bpdf, candf, info);
    if (phase != PHASE_PENDING)
        return;
    int32_t cand = __float_as_int(candf);
    pt::Xorshift32 rng{rng_state};
    const Surface s = surface_current(d, pix);
    const pt::Material& m = d.mats[s.mat];
    const int32_t nl = r.s.number_of_initial_light_candidates;
    const int32_t nb = r.s.number_of_initial_bsdf_candidates;
    const float4 ra = r.rays[2 * (size_t)pix], rb = r.rays[2 * (size_t)pix + 1];
    const f3 o = ld3(ra), dir = ld3(rb);
    const int4 h = r.hits[pix];
    if (h.x >= 0)
    {
        const int32_t hp = h.x;
        const float t = __int_as_float(h.y);
        const Color emission = d.mats[d.mat_idx[hp]].emission;
        if (!pt::is_black(emission))
        {
            const f3 point = o + t * dir;
            const uint32_t flags = (pt::RESTIR_FLAG_UNOCCLUDED | info) & 0xffu;
            f3 tdir;
            float tdist;
            const float target = target_function(d, s, point, hp, flags, tdir, tdist);
            const f3 hgn = pt::normalize(pt::normalize(tri_normal_unnormalized(d, hp)));
            const float cos_l = cos_at_light(hgn, -1.0f * dir);
            float lpdf = 0.0f;
            // with presampling, the light sampler is the presampling's
            // (pdf_of_emissive_triangle_hit_solid_angle<ReSTIR_DI_
            // LightPresamplingStrategy>, :360-366): power
            const int32_t strategy = r.presampled ? MIPT_LSS_BASE_POWER : d.light_base;
            if (!(pt::dot(dir, s.sn) < 0.0f))
                lpdf = area_to_solid_angle(emissive_hit_area_pdf_with(d, strategy, d.areas[hp], emission), t, cos_l);
            if (!(bpdf <= 0.0f))
            {
                lpdf *= (1.0f - env_candidate_probability(d, r));
                const float mis = balance_heuristic(bpdf, (float)nb, lpdf, (float)nl);
                float pdf_area = bpdf;
                pdf_area /= (t * t);
                pdf_area *= cos_l;
                float weight = mis * target / pdf_area;
                if (r.s.guard_backfacing_bsdf_candidates && !(pdf_area > 0.0f))
                    weight = 0.0f;
                add_candidate(res, point, hp, target, flags, weight, rng);
            }
        }
    }
    else if (env_sampled(d.env) && pt::dot(s.sn, dir) > 0.0f)
    {
        // an envmap hit becomes an envmap sample (:396-436); its light pdf is
        // the envmap sampler's, refractions excluded
        float epdf;
        env_eval(d.env, dir, epdf);
        const f3 point = env_rotate(d.env.to_env, dir);
        const uint32_t flags = (pt::RESTIR_FLAG_UNOCCLUDED | pt::RESTIR_FLAG_ENVMAP | info) & 0xffu;
        f3 tdir;
        float tdist;
        const float target = target_function(d, s, point, -1, flags, tdir, tdist);
        if (!(bpdf <= 0.0f))
        {
            epdf *= env_candidate_probability(d, r);
            const float mis = balance_heuristic(bpdf, (float)nb, epdf, (float)nl);
            const float weight = mis * target / bpdf;
            add_candidate(res, point, -1, target, flags, weight, rng);
        }
    }
    cand++;
    if (next_bsdf_candidate(d, r, pix, s, m, cand, rng, bpdf, info))
    {
        state_store(r, pix, res, rng.state, PHASE_PENDING, bpdf, __int_as_float(cand), info);
        return;
    }
    write_inactive(r.rays, pix);
    uint32_t need_vis;
    finish_initial(d, r, pix, s, res, need_vis);
    state_store(r, pix, res, rng.state, PHASE_DONE, 0.0f, 0.0f, need_vis);
}

__global__ void k_ic_finish(RsData r)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= r.n)
        return;
    Res res;
    uint32_t rng_state, need_vis;
    int32_t phase;
    float x0, x1;
    state_load(r, pix, res, rng_state, phase, x0, x1, need_vis);
    if (phase != PHASE_DONE)
        return;
    if (need_vis)
    {
        if (r.occ[(size_t)r.n + pix])
            res.ucw = -1.0f;
        else
            res.flags |= pt::RESTIR_FLAG_UNOCCLUDED;
    }
    res_store(r.out, pix, res);
}

// check_neighbor_similarity_heuristics(..., previous_frame = true)
// (NeighborSimilarity.h:52-107): primary hit and position of the previous
// frame's G-buffer, the neighbour's normal from the current one.
__device__ __forceinline__ bool similar_prev(const PtData& d, const RsData& r, int32_t npix, uint32_t cpix, f3 cpoint,
                                             f3 cnormal)
{
    if ((uint32_t)npix == cpix)
        return true;
    if (r.pg_prim[npix] == -1)
        return false;
    const f3 np = ld3(r.pg_pos[npix]);
    const f3 nn = r.s.reject_using_geometric_normals ? gb_normal(d.gb_gn[npix]) : gb_normal(d.gb_sn[npix]);
    bool plane = true, normal = true;
    if (r.s.use_plane_distance_heuristic)
        plane = fabsf(pt::dot(np - cpoint, cnormal)) < r.s.plane_distance_threshold;
    if (r.s.use_normal_similarity_heuristic)
        normal = pt::dot(cnormal, nn) > r.s.normal_similarity_angle_precomp;
    return plane && normal;
}

// A pixel coordinate from a rounded float: values out of any viewport's
// range go to a coordinate far outside (their int conversion is undefined in
// the reference).
__device__ __forceinline__ int32_t pixel_coord(float v) { return (v > -1.0e9f && v < 1.0e9f) ? (int32_t)v : -1000000000; }

// find_temporal_neighbor_index (UtilsTemporal.h:37-99): (index, x, y) of the
// temporal neighbour. Accumulating, the pixel itself. Else the primary hit
// reprojected by the previous frame's view-projection (matrix_X_point,
// Math.h:617-633), rounded, optionally permuted within 4 x 4 blocks
// (apply_permutation_sampling, :11-22), then up to max_neighbor_search_count
// random pixels within neighbor_search_radius (x offset drawn first), the
// first that passes the similarity heuristics against the previous G-buffer;
// index -1 without one. (x, y) is the rounded reprojected pixel.
__device__ int3 temporal_neighbor(const PtData& d, const RsData& r, uint32_t pix, f3 cnormal, pt::Xorshift32& rng)
{
    const int32_t W = d.width, H = d.height;
    if (r.accumulate)
        return make_int3((int32_t)pix, (int32_t)(pix % (uint32_t)W), (int32_t)(pix / (uint32_t)W));
    const f3 p = ld3(d.gb_pos[pix]);
    const float(*m)[4] = r.prev_vp;
    const float xt = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
    const float yt = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
    const float wt = m[3][0] * p.x + m[3][1] * p.y + m[3][2] * p.z + m[3][3];
    float inv_w = 1.0f;
    if (!(wt < 1.0e-10f && wt > -1.0e-10f))
        inv_w = 1.0f / wt;
    float sx = xt * inv_w, sy = yt * inv_w;
    sx += 1.0f;
    sy += 1.0f;
    sx *= 0.5f;
    sy *= 0.5f;
    float fx = sx * (float)W, fy = sy * (float)H;
    fx -= 0.5f;
    fy -= 0.5f;
    int32_t idx = -1;
    for (int32_t i = 0; i < r.s.max_neighbor_search_count + 1; i++)
    {
        float ox = 0.0f, oy = 0.0f;
        if (i > 0)
        {
            const float rx = rng() - 0.5f;
            const float ry = rng() - 0.5f;
            ox = rx * (float)r.s.neighbor_search_radius;
            oy = ry * (float)r.s.neighbor_search_radius;
        }
        int32_t x = pixel_coord(roundf(fx + ox)), y = pixel_coord(roundf(fy + oy));
        if (r.s.use_permutation_sampling && i == 0)
        {
            const int32_t bits = r.s.permutation_sampling_random_bits;
            const int32_t dx = bits & 3, dy = (bits >> 2) & 3;
            x += dx;
            y += dy;
            x ^= 3;
            y ^= 3;
            x -= dx;
            y -= dy;
        }
        if (x < 0 || x >= W || y < 0 || y >= H)
            continue;
        idx = x + y * W;
        if (similar_prev(d, r, idx, pix, p, cnormal))
            break;
        idx = -1;
    }
    return make_int3(idx, pixel_coord(roundf(fx)), pixel_coord(roundf(fy)));
}

// ---------------------------------------------------------------------------
// TEMPORAL REUSE (TemporalReuse.h:50-296). The temporal neighbour is the
// pixel itself when accumulating, the reprojected pixel otherwise; k_tp_start
// keeps its index in the pass state (slot 3.z) with the pixel's random state
// after the neighbour search, which k_tp_finish resumes.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void state_store_neighbor(const RsData& r, uint32_t pix, int32_t neighbor)
{
    r.state[4 * (size_t)pix + 3].z = __int_as_float(neighbor);
}

__global__ void k_tp_start(PtData d, RsData r)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= r.n)
        return;
    write_inactive(r.rays, pix);
    write_inactive(r.rays, (size_t)r.n + pix);
    if (d.gb_prim[pix] == -1)
    {
        state_store(r, pix, res_empty(), 0u, PHASE_SKIP, 0.0f, 0.0f, 0u);
        return;
    }
    pt::Xorshift32 rng{pt::pixel_seed(pix, d.sample_number, d.random_number, d.freeze != 0)};
    const Surface center = surface_current(d, pix);
    const int32_t tix = temporal_neighbor(d, r, pix, center.sn, rng).x;
    if (tix == -1 || d.freeze)
    {
        state_store(r, pix, res_empty(), 0u, PHASE_COPY, 0.0f, 0.0f, 0u);
        return;
    }
    const Res tn = res_load(r.in, (uint32_t)tix);
    if (tn.M == 0)
    {
        state_store(r, pix, res_empty(), 0u, PHASE_COPY, 0.0f, 0.0f, 0u);
        return;
    }
    const Surface prev = surface_previous(d, r, (uint32_t)tix);
    if (is_emissive(d, prev.mat))
    {
        state_store(r, pix, res_empty(), 0u, PHASE_COPY, 0.0f, 0.0f, 0u);
        return;
    }
    const Res init = res_load(r.init, pix);
    uint32_t rays = 0u;
    float tf_center = 0.0f;
    if (tn.ucw > 0.0f)
    {
        f3 dir;
        float dist;
        tf_center = target_function(d, center, tn.point, tn.tri, tn.flags, dir, dist);
        if (tf_center != 0.0f)
        {
            write_ray(r.rays, pix, center.p, center.prim, dir, dist - 1.0e-4f);
            rays |= 1u;
        }
    }
    f3 dir;
    float dist;
    const float tf_csn = target_function(d, prev, init.point, init.tri, init.flags, dir, dist);
    if (tf_csn != 0.0f)
    {
        write_ray(r.rays, (size_t)r.n + pix, prev.p, prev.prim, dir, dist - 1.0e-4f);
        rays |= 2u;
    }
    state_store(r, pix, res_empty(), rng.state, PHASE_DONE, tf_center, tf_csn, rays);
    state_store_neighbor(r, pix, tix);
}

// TemporalReuse.h:100-296 with the pairwise MIS weights of
// TemporalMISWeight.h:164-254.
__global__ void k_tp_finish(PtData d, RsData r)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= r.n)
        return;
    Res unused;
    uint32_t rng_state, rays;
    int32_t phase;
    float tf_center, tf_csn;
    state_load(r, pix, unused, rng_state, phase, tf_center, tf_csn, rays);
    if (phase == PHASE_SKIP)
        return;
    const Res init = res_load(r.init, pix);
    if (phase == PHASE_COPY)
    {
        res_store(r.out, pix, init);
        return;
    }
    const int32_t tix = __float_as_int(r.state[4 * (size_t)pix + 3].z);
    const Res tn = res_load(r.in, (uint32_t)tix);
    if (rays & 1u)
        tf_center = apply_visibility(tf_center, r.occ[pix] != 0);
    if (rays & 2u)
        tf_csn = apply_visibility(tf_csn, r.occ[(size_t)r.n + pix] != 0);
    pt::Xorshift32 rng{rng_state};
    Res out = res_empty();
    const bool cw = r.s.use_confidence_weights != 0;
    const float tn_M = cw ? (float)tn.M : 1.0f;
    const float center_M = cw ? (float)init.M : 1.0f;
    const float conf_sum = cw ? (float)tn.M : 1.0f;
    // temporal neighbour
    const float tf_at_neighbor = tn.target;
    const float nume = tf_at_neighbor * tn_M;
    const float denom = tf_at_neighbor * conf_sum + tf_center * center_M;
    const float mi = denom == 0.0f ? 0.0f : nume / denom;
    const float tf_csc = init.target;
    const float nume_mc = tf_csc * center_M;
    const float denom_mc = tf_csn * conf_sum + tf_csc * center_M;
    const float conf_mult = cw ? tn_M / conf_sum : 1.0f;
    float mc = 0.0f;
    if (denom_mc != 0.0f)
        mc += nume_mc / denom_mc * conf_mult;
    if (combine_with(out, tn, mi, tf_center, 1.0f, rng))
        out.flags |= pt::RESTIR_FLAG_UNOCCLUDED;
    // initial candidates
    const float mis_init = mc == 0.0f ? 1.0f : mc;
    if (combine_with(out, init, mis_init, init.target, 1.0f, rng))
        out.flags |= pt::RESTIR_FLAG_UNOCCLUDED;
    res_end_normalized(out, 1.0f, 1.0f);
    if (r.s.m_cap > 0)
        out.M = out.M < r.s.m_cap ? out.M : r.s.m_cap;
    res_store(r.out, pix, out);
}

// ---------------------------------------------------------------------------
// SPATIAL REUSE
// ---------------------------------------------------------------------------

// get_spatial_neighbor_pixel_index (UtilsSpatial.h:177-259), neighbour
// `i` < K, uniform disk sample (Sampling.h:50-57); -1 outside the image.
__device__ __forceinline__ int32_t spatial_neighbor(const PtData& d, const RsData& r, int32_t cx, int32_t cy,
                                                    pt::Xorshift32& nrng)
{
    const float u = nrng();
    const float v = nrng();
    const float rad = (float)r.s.spatial_reuse_radius * sqrtf(v);
    float sn_, cs_;
    pt::sincos_fixed(6.28318530717958647693f * u, sn_, cs_);
    const float ox = rad * cs_, oy = rad * sn_;
    const int32_t nx = cx + (int32_t)roundf(ox);
    const int32_t ny = cy + (int32_t)roundf(oy);
    if (nx < 0 || nx >= d.width || ny < 0 || ny >= d.height)
        return -1;
    return nx + ny * d.width;
}

// check_neighbor_similarity_heuristics (NeighborSimilarity.h:46-107), current frame.
__device__ __forceinline__ bool similar(const PtData& d, const RsData& r, int32_t npix, uint32_t cpix, f3 cpoint,
                                        f3 cnormal)
{
    if ((uint32_t)npix == cpix)
        return true;
    if (d.gb_prim[npix] == -1)
        return false;
    const f3 np = ld3(d.gb_pos[npix]);
    const f3 nn = r.s.reject_using_geometric_normals ? gb_normal(d.gb_gn[npix]) : gb_normal(d.gb_sn[npix]);
    bool plane = true, normal = true;
    if (r.s.use_plane_distance_heuristic)
        plane = fabsf(pt::dot(np - cpoint, cnormal)) < r.s.plane_distance_threshold;
    if (r.s.use_normal_similarity_heuristic)
        normal = pt::dot(cnormal, nn) > r.s.normal_similarity_angle_precomp;
    return plane && normal;
}

__global__ __launch_bounds__(256, 2) void k_sp_start(PtData d, RsData r)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= r.n)
        return;
    const int32_t K = r.s.spatial_reuse_neighbor_count;
    for (int32_t k = 0; k < K; k++)
        write_inactive(r.rays, (size_t)k * r.n + pix);
    if (d.gb_prim[pix] == -1)
    {
        state_store(r, pix, res_empty(), 0u, PHASE_SKIP, 0.0f, 0.0f, 0u);
        return;
    }
    const int32_t cx = (int32_t)(pix % (uint32_t)d.width), cy = (int32_t)(pix / (uint32_t)d.width);
    pt::Xorshift32 rng{pt::pixel_seed(pix, d.sample_number, d.random_number, d.freeze != 0)};
    const Res center = res_load(r.in, pix);
    const Surface cs = surface_current(d, pix);
    // setup_adaptive_directional_spatial_reuse (UtilsSpatial.h:17-31): the
    // neighbour stream is seeded from this pixel's stream.
    const uint32_t nseed = rng.next();
    const f3 cnormal = r.s.reject_using_geometric_normals ? cs.gn : cs.sn;
    // count_valid_spatial_neighbors (UtilsSpatial.h:302-339)
    int32_t valid_count = 0, valid_M = 0;
    uint32_t cache = 0u;
    {
        pt::Xorshift32 nrng{nseed};
        for (int32_t i = 0; i < K; i++)
        {
            const int32_t np = spatial_neighbor(d, r, cx, cy, nrng);
            if (np == -1)
                continue;
            if (!similar(d, r, np, pix, cs.p, cnormal))
                continue;
            valid_M += __float_as_int(r.in[2 * (size_t)np + 1].x);
            valid_count++;
            cache |= 1u << i;
        }
    }
    const bool cw = r.s.use_confidence_weights != 0;
    const float center_M = cw ? (float)center.M : 1.0f;
    const float conf_sum = cw ? (float)valid_M : 1.0f;
    const float vdt = cw ? 1.0f : (float)valid_count;
    const bool update_mc = center.M > 0 && center.ucw > 0.0f;
    const float tf_csc = center.target;
    const float nume_mc = tf_csc / vdt * center_M;
    float* aux = r.aux + 2 * (size_t)K * pix;
    for (int32_t i = 0; i < K; i++)
    {
        aux[2 * i] = 0.0f;
        aux[2 * i + 1] = -1.0f;
    }
    Res out = res_empty();
    pt::Xorshift32 nrng{nseed};
    const int32_t start = valid_M == 0 ? K : 0;
    for (int32_t i = start; i < K; i++)
    {
        if (!(cache & (1u << i)))
        {
            nrng.next();
            nrng.next();
            continue;
        }
        const int32_t np = spatial_neighbor(d, r, cx, cy, nrng);
        if (np == -1)
            continue;
        const Res nres = res_load(r.in, np);
        float tf_center = 0.0f;
        if (nres.ucw > 0.0f)
        {
            f3 dir;
            float dist;
            tf_center = target_function(d, cs, nres.point, nres.tri, nres.flags, dir, dist);
        }
        // ReSTIRSpatialResamplingMISWeight<PAIRWISE_MIS> (SpatialMISWeight.h:118-234), neighbour
        const float res_M = cw ? (float)nres.M : 1.0f;
        const float nume = nres.target * res_M;
        const float denom = nres.target * conf_sum + tf_center / vdt * center_M;
        const float mi = denom == 0.0f ? 0.0f : nume / denom;
        if (update_mc)
        {
            const Surface ns = surface_current(d, (uint32_t)np);
            f3 dir;
            float dist;
            const float tf_csn = target_function(d, ns, center.point, center.tri, center.flags, dir, dist);
            const float conf_mult = cw ? (conf_sum == 0.0f ? 0.0f : res_M / conf_sum) : 1.0f;
            aux[2 * i] = tf_csn;
            aux[2 * i + 1] = conf_mult;
            if (tf_csn != 0.0f)
                write_ray(r.rays, (size_t)i * r.n + pix, ns.p, ns.prim, dir, dist - 1.0e-4f);
        }
        if (combine_with(out, nres, mi / vdt, tf_center, 1.0f, rng))
            out.flags &= ~pt::RESTIR_FLAG_UNOCCLUDED;
    }
    state_store(r, pix, out, rng.state, PHASE_PENDING, nume_mc, conf_sum, (uint32_t)valid_count);
}

__global__ void k_sp_center(PtData d, RsData r)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= r.n)
        return;
    write_inactive(r.rays, pix);
    Res out;
    uint32_t rng_state, valid_count;
    int32_t phase;
    float nume_mc, conf_sum;
    state_load(r, pix, out, rng_state, phase, nume_mc, conf_sum, valid_count);
    if (phase == PHASE_SKIP)
        return;
    const int32_t K = r.s.spatial_reuse_neighbor_count;
    // valid_neighbor_division_term (SpatialMISWeight.h:150): 1 with
    // confidence weights, else the number of valid neighbours.
    const float vdt = r.s.use_confidence_weights ? 1.0f : (float)valid_count;
    const float* aux = r.aux + 2 * (size_t)K * pix;
    float mc = 0.0f;
    for (int32_t i = 0; i < K; i++)
    {
        const float conf_mult = aux[2 * i + 1];
        if (conf_mult < 0.0f)
            continue;
        float tf_csn = aux[2 * i];
        if (tf_csn != 0.0f)
            tf_csn = apply_visibility(tf_csn, r.occ[(size_t)i * r.n + pix] != 0);
        const float denom_mc = tf_csn * conf_sum + nume_mc;
        if (denom_mc != 0.0f)
            mc += nume_mc / denom_mc / vdt * conf_mult;
    }
    pt::Xorshift32 rng{rng_state};
    const Res center = res_load(r.in, pix);
    const float tf_center = center.ucw > 0.0f ? center.target : 0.0f;
    const float mis = mc == 0.0f ? 1.0f : mc;
    if (combine_with(out, center, mis, tf_center, 1.0f, rng))
        out.flags |= center.flags & pt::RESTIR_FLAG_UNOCCLUDED;
    res_end_normalized(out, 1.0f, 1.0f);
    uint32_t need_vis = 0u;
    if (r.kill_after && !(out.ucw <= 0.0f && out.wsum <= 0.0f) && !(out.flags & pt::RESTIR_FLAG_UNOCCLUDED))
    {
        const f3 p = ld3(d.gb_pos[pix]);
        float dist;
        const f3 dir = sample_direction(d, p, out.point, out.flags, dist);
        write_ray(r.rays, pix, p, d.gb_prim[pix], dir, dist - 1.0e-4f);
        need_vis = 1u;
    }
    state_store(r, pix, out, rng.state, PHASE_DONE, 0.0f, 0.0f, need_vis);
}

__global__ void k_sp_finish(RsData r)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= r.n)
        return;
    Res out;
    uint32_t rng_state, need_vis;
    int32_t phase;
    float x0, x1;
    state_load(r, pix, out, rng_state, phase, x0, x1, need_vis);
    if (phase != PHASE_DONE)
        return;
    if (need_vis)
    {
        if (r.occ[pix])
            out.ucw = -1.0f;
        else
            out.flags |= pt::RESTIR_FLAG_UNOCCLUDED;
    }
    if (r.s.m_cap > 0)
        out.M = out.M < r.s.m_cap ? out.M : r.s.m_cap;
    res_store(r.out, pix, out);
}

// ---------------------------------------------------------------------------
// FUSED SPATIOTEMPORAL REUSE (FusedSpatiotemporalReuse.h:152-566): the
// temporal neighbour is found by find_temporal_neighbor_index (the pixel
// itself when accumulating) and the spatial neighbours are drawn around its
// reprojected position. Temporal neighbour and spatial neighbours
// come from last frame's output reservoirs; neighbour surfaces and the
// similarity tests read the previous frame's G-buffer
// (use_prev_frame_g_buffer(), RenderSettings.h:263-272). Pairwise MIS
// (SpatiotemporalMISWeight.h:139-224) with the temporal neighbour resampled
// first, the initial candidates last.
//
// Three kernels around two ray launches: k_fst_start evaluates every target
// function that needs visibility (the temporal sample at the centre; the
// centre's initial sample at the temporal and at each spatial neighbour's
// surface) and writes their shadow rays; k_fst_combine applies the
// visibility and resamples in the reference's order, then writes the
// visibility-reuse ray; k_fst_finish kills or keeps the sample.
// Per-pixel aux layout (stride fst_aux_stride(K)): [0] temporal target at the
// centre (unoccluded), [1] neighbour seed, [2] valid count, [3] valid M sum,
// [4] heuristics cache, [5] ray mask (bit 0 centre, bit 1 temporal neighbour),
// [6], [7] the reprojected pixel (the spatial neighbours' centre), then
// (target, confidence multiplier) for the temporal neighbour and for each
// spatial neighbour (multiplier < 0: no entry). The temporal neighbour's
// index is in the pass state (slot 3.z).
// ---------------------------------------------------------------------------
__device__ __forceinline__ size_t fst_aux_stride(int32_t K) { return 8 + 2 * (size_t)(K + 1); }

// The fused pass seeds with wang_hash(((pix + 1) * (sample + 1)) ^ random)
// (FusedSpatiotemporalReuse.h:173), an XOR where the other passes multiply.
__device__ __forceinline__ uint32_t fused_seed(uint32_t pix, uint32_t sample_number, uint32_t random_number, bool freeze)
{
    return freeze ? pt::wang_hash(pix + 1u) : pt::wang_hash(((pix + 1u) * (sample_number + 1u)) ^ random_number);
}

__global__ __launch_bounds__(256, 2) void k_fst_start(PtData d, RsData r)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= r.n)
        return;
    const int32_t K = r.s.spatial_reuse_neighbor_count;
    for (int32_t k = 0; k < K + 2; k++)
        write_inactive(r.rays, (size_t)k * r.n + pix);
    if (d.gb_prim[pix] == -1)
    {
        state_store(r, pix, res_empty(), 0u, PHASE_SKIP, 0.0f, 0.0f, 0u);
        return;
    }
    pt::Xorshift32 rng{fused_seed(pix, d.sample_number, d.random_number, d.freeze != 0)};
    const uint32_t nseed = rng.next();
    const Surface cs = surface_current(d, pix);
    // load_spatiotemporal_neighbor_data (:71-101): no temporal neighbour or
    // frozen random leaves the temporal reservoir empty; the spatial
    // neighbours are drawn around the reprojected pixel
    const int3 tnb = temporal_neighbor(d, r, pix, cs.sn, rng);
    const Res tn = (tnb.x == -1 || d.freeze) ? res_empty() : res_load(r.in, (uint32_t)tnb.x);
    const int32_t cx = tnb.y, cy = tnb.z;
    const f3 cnormal = r.s.reject_using_geometric_normals ? cs.gn : cs.sn;
    // count_valid_spatiotemporal_neighbors (:103-148)
    int32_t valid_count = 0, valid_M = 0;
    uint32_t cache = 0u;
    {
        pt::Xorshift32 nrng{nseed};
        for (int32_t i = 0; i < K; i++)
        {
            const int32_t np = spatial_neighbor(d, r, cx, cy, nrng);
            if (np == -1)
                continue;
            if (!similar_prev(d, r, np, pix, cs.p, cnormal))
                continue;
            valid_M += __float_as_int(r.in[2 * (size_t)np + 1].x);
            valid_count++;
            cache |= 1u << i;
        }
    }
    if (tn.M > 0)
    {
        valid_count++;
        valid_M += tn.M;
    }
    const Res init = res_load(r.init, pix);
    const bool cw = r.s.use_confidence_weights != 0;
    const float conf_sum = cw ? (float)valid_M : 1.0f;
    const bool update_mc = init.M > 0 && init.ucw > 0.0f;
    float* aux = r.aux + fst_aux_stride(K) * pix;
    for (int32_t i = 0; i < K + 1; i++)
    {
        aux[8 + 2 * i] = 0.0f;
        aux[8 + 2 * i + 1] = -1.0f;
    }
    uint32_t rays = 0u;
    float tf_center_t = 0.0f;
    if (tn.M > 0)
    {
        f3 dir;
        float dist;
        if (tn.ucw > 0.0f)
        {
            tf_center_t = target_function(d, cs, tn.point, tn.tri, tn.flags, dir, dist);
            if (tf_center_t != 0.0f)
            {
                write_ray(r.rays, pix, cs.p, cs.prim, dir, dist - 1.0e-4f);
                rays |= 1u;
            }
        }
        if (update_mc)
        {
            const Surface ts = surface_previous(d, r, (uint32_t)tnb.x);
            const float tf_csn = target_function(d, ts, init.point, init.tri, init.flags, dir, dist);
            const float tn_M = cw ? (float)tn.M : 1.0f;
            aux[8] = tf_csn;
            aux[9] = cw ? tn_M / conf_sum : 1.0f;
            if (tf_csn != 0.0f)
            {
                write_ray(r.rays, (size_t)r.n + pix, ts.p, ts.prim, dir, dist - 1.0e-4f);
                rays |= 2u;
            }
        }
    }
    if (update_mc)
    {
        pt::Xorshift32 nrng{nseed};
        const int32_t start = valid_M == 0 ? K : 0;
        for (int32_t i = start; i < K; i++)
        {
            if (!(cache & (1u << i)))
            {
                nrng.next();
                nrng.next();
                continue;
            }
            const int32_t np = spatial_neighbor(d, r, cx, cy, nrng);
            if (np == -1)
                continue;
            const Surface ns = surface_previous(d, r, (uint32_t)np);
            f3 dir;
            float dist;
            const float tf_csn = target_function(d, ns, init.point, init.tri, init.flags, dir, dist);
            const float res_M = cw ? (float)__float_as_int(r.in[2 * (size_t)np + 1].x) : 1.0f;
            aux[8 + 2 * (i + 1)] = tf_csn;
            aux[8 + 2 * (i + 1) + 1] = cw ? res_M / conf_sum : 1.0f;
            if (tf_csn != 0.0f)
                write_ray(r.rays, (size_t)(i + 2) * r.n + pix, ns.p, ns.prim, dir, dist - 1.0e-4f);
        }
    }
    aux[0] = tf_center_t;
    aux[1] = __uint_as_float(nseed);
    aux[2] = __int_as_float(valid_count);
    aux[3] = __int_as_float(valid_M);
    aux[4] = __uint_as_float(cache);
    aux[5] = __uint_as_float(rays);
    aux[6] = __int_as_float(cx);
    aux[7] = __int_as_float(cy);
    state_store(r, pix, res_empty(), rng.state, PHASE_PENDING, 0.0f, 0.0f, 0u);
    state_store_neighbor(r, pix, tnb.x);
}

__global__ __launch_bounds__(256, 2) void k_fst_combine(PtData d, RsData r)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= r.n)
        return;
    Res unused;
    uint32_t rng_state, x2;
    int32_t phase;
    float x0, x1;
    state_load(r, pix, unused, rng_state, phase, x0, x1, x2);
    write_inactive(r.rays, pix);
    if (phase == PHASE_SKIP)
        return;
    const int32_t K = r.s.spatial_reuse_neighbor_count;
    const float* aux = r.aux + fst_aux_stride(K) * pix;
    const int32_t cx = __float_as_int(aux[6]), cy = __float_as_int(aux[7]);
    const int32_t tix = __float_as_int(r.state[4 * (size_t)pix + 3].z);
    const uint32_t nseed = __float_as_uint(aux[1]);
    const int32_t valid_count = __float_as_int(aux[2]);
    const int32_t valid_M = __float_as_int(aux[3]);
    const uint32_t cache = __float_as_uint(aux[4]);
    const uint32_t rays = __float_as_uint(aux[5]);
    pt::Xorshift32 rng{rng_state};
    const Surface cs = surface_current(d, pix);
    const Res tn = (tix == -1 || d.freeze) ? res_empty() : res_load(r.in, (uint32_t)tix);
    const Res init = res_load(r.init, pix);
    const bool cw = r.s.use_confidence_weights != 0;
    const float center_M = cw ? (float)init.M : 1.0f;
    const float conf_sum = cw ? (float)valid_M : 1.0f;
    const float vdt = cw ? 1.0f : (float)valid_count;
    const bool update_mc = init.M > 0 && init.ucw > 0.0f;
    const float nume_mc = init.target / vdt * center_M;
    Res out = res_empty();
    float mc = 0.0f;
    if (tn.M > 0)
    {
        float tf_center = aux[0];
        if (rays & 1u)
            tf_center = apply_visibility(tf_center, r.occ[pix] != 0);
        const float tn_M = cw ? (float)tn.M : 1.0f;
        const float nume = tn.target * tn_M;
        const float denom = tn.target * conf_sum + tf_center / vdt * center_M;
        const float mi = denom == 0.0f ? 0.0f : nume / denom;
        if (update_mc)
        {
            float tf_csn = aux[8];
            if (rays & 2u)
                tf_csn = apply_visibility(tf_csn, r.occ[(size_t)r.n + pix] != 0);
            const float denom_mc = tf_csn * conf_sum + nume_mc;
            if (denom_mc != 0.0f)
                mc += nume_mc / denom_mc / vdt * aux[9];
        }
        if (combine_with(out, tn, mi / vdt, tf_center, 1.0f, rng))
            out.flags |= pt::RESTIR_FLAG_UNOCCLUDED;
    }
    pt::Xorshift32 nrng{nseed};
    const int32_t start = valid_M == 0 ? K : 0;
    for (int32_t i = start; i < K + 1; i++)
    {
        if (i < K && !(cache & (1u << i)))
        {
            nrng.next();
            nrng.next();
            continue;
        }
        const int32_t np = i == K ? (int32_t)pix : spatial_neighbor(d, r, cx, cy, nrng);
        if (np == -1)
            continue;
        const Res nres = i == K ? init : res_load(r.in, (uint32_t)np);
        float tf_center = 0.0f;
        if (nres.ucw > 0.0f)
        {
            if (i == K)
                tf_center = nres.target;
            else
            {
                f3 dir;
                float dist;
                tf_center = target_function(d, cs, nres.point, nres.tri, nres.flags, dir, dist);
            }
        }
        float w;
        if (i < K)
        {
            const float res_M = cw ? (float)nres.M : 1.0f;
            const float nume = nres.target * res_M;
            const float denom = nres.target * conf_sum + tf_center / vdt * center_M;
            const float mi = denom == 0.0f ? 0.0f : nume / denom;
            if (update_mc)
            {
                float tf_csn = aux[8 + 2 * (i + 1)];
                if (tf_csn != 0.0f)
                    tf_csn = apply_visibility(tf_csn, r.occ[(size_t)(i + 2) * r.n + pix] != 0);
                const float denom_mc = tf_csn * conf_sum + nume_mc;
                if (denom_mc != 0.0f)
                    mc += nume_mc / denom_mc / vdt * aux[8 + 2 * (i + 1) + 1];
            }
            w = mi / vdt;
        }
        else
            w = mc == 0.0f ? 1.0f : mc;
        if (combine_with(out, nres, w, tf_center, 1.0f, rng))
        {
            if (i == K)
                out.flags |= nres.flags & pt::RESTIR_FLAG_UNOCCLUDED;
            else
                out.flags &= ~pt::RESTIR_FLAG_UNOCCLUDED;
        }
    }
    res_end_normalized(out, 1.0f, 1.0f);
    uint32_t need_vis = 0u;
    if (r.kill_after && !(out.ucw <= 0.0f && out.wsum <= 0.0f) && !(out.flags & pt::RESTIR_FLAG_UNOCCLUDED))
    {
        const f3 p = ld3(d.gb_pos[pix]);
        float dist;
        const f3 dir = sample_direction(d, p, out.point, out.flags, dist);
        write_ray(r.rays, pix, p, d.gb_prim[pix], dir, dist - 1.0e-4f);
        need_vis = 1u;
    }
    state_store(r, pix, out, rng.state, PHASE_DONE, 0.0f, 0.0f, need_vis);
}

__global__ void k_reservoir_clear(float4* buf, uint32_t n)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix < n)
        res_store(buf, pix, res_empty());
}

inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

RsData make_rs_data(Ctx& ctx)
{
    RenderState& R = ctx.render;
    RsData r{};
    r.state = R.rs_state.as<float4>();
    r.rays = R.rs_rays.as<float4>();
    r.occ = R.rs_occ.as<uint8_t>();
    r.hits = R.rs_hits.as<int4>();
    r.aux = R.rs_aux.as<float>();
    r.pg_prim = R.pg_prim.as<int32_t>();
    r.pg_pos = R.pg_pos.as<float4>();
    r.pg_gn = R.pg_gn.as<uint32_t>();
    r.pg_sn = R.pg_sn.as<uint32_t>();
    r.pg_mat = R.pg_mat.as<int32_t>();
    r.n = R.pixels;
    r.s = R.restir;
    r.accumulate = R.settings.accumulate;
    std::memcpy(r.prev_vp, R.prev_camera.view_projection, sizeof(r.prev_vp));
    std::memcpy(r.prev_cam_pos, R.prev_camera.position, sizeof(r.prev_cam_pos));
    return r;
}

// configure_output_buffer (ReSTIRDIRenderPass.cpp:659-677) after the last
// ReSTIR pass of a frame.
void end_frame(RenderState& R, int32_t output)
{
    R.restir_shade_buffer = output;
    R.last_output = output;
    R.odd_frame = !R.odd_frame;
}

void check_ready(Ctx& ctx)
{
    RenderState& R = ctx.render;
    if (R.settings.direct_light_sampling_strategy != MIPT_LSS_RESTIR_DI)
        fail(MIPT_ERR_INVALID_ARGUMENT, "ReSTIR DI passes need direct_light_sampling_strategy = MIPT_LSS_RESTIR_DI");
    if (!R.reservoirs[0].ptr)
        fail(MIPT_ERR_INVALID_ARGUMENT, "ReSTIR DI buffers are not allocated");
}

} // namespace

bool restir_needs_prev_gbuffer(const RenderState& R)
{
    return R.settings.direct_light_sampling_strategy == MIPT_LSS_RESTIR_DI && R.restir.do_temporal_reuse_pass;
}

float4* restir_reservoir_buffer(Ctx& ctx, int32_t which)
{
    return ctx.render.reservoirs[which].as<float4>();
}

void allocate_restir_buffers(Ctx& ctx)
{
    RenderState& R = ctx.render;
    const size_t n = R.pixels;
    if (R.settings.direct_light_sampling_strategy != MIPT_LSS_RESTIR_DI || n == 0)
        return;
    const int32_t K = R.restir.spatial_reuse_neighbor_count;
    // ray slots per pixel: K spatial, 2 temporal, K + 2 fused
    const size_t slots = (size_t)K + 2;
    for (int i = 0; i < 3; i++)
        R.reservoirs[i].reserve(32 * n);
    if (restir_needs_prev_gbuffer(R))
    {
        R.pg_prim.reserve(4 * n);
        R.pg_pos.reserve(16 * n);
        R.pg_gn.reserve(4 * n);
        R.pg_sn.reserve(4 * n);
        R.pg_mat.reserve(4 * n);
    }
    R.rs_state.reserve(64 * n);
    R.rs_rays.reserve(32 * n * slots);
    R.rs_occ.reserve(n * slots);
    R.rs_hits.reserve(16 * n);
    R.rs_aux.reserve(4 * n * std::max<size_t>(2 * (size_t)(K > 0 ? K : 1), 8 + 2 * ((size_t)K + 1)));
    if (R.restir.do_light_presampling)
    {
        R.presampled.reserve(48 * (size_t)R.restir.light_presampling_number_of_subsets *
                             (size_t)R.restir.light_presampling_subset_size);
        R.presampled_ready = false;
    }
}

void restir_reset(Ctx& ctx, hipStream_t s)
{
    RenderState& R = ctx.render;
    if (R.settings.direct_light_sampling_strategy != MIPT_LSS_RESTIR_DI || !R.reservoirs[0].ptr)
        return;
    for (int i = 0; i < 3; i++)
        hipLaunchKernelGGL(k_reservoir_clear, dim3(blocks_for(R.pixels, 256)), dim3(256), 0, s,
                           R.reservoirs[i].as<float4>(), R.pixels);
    HIP_CHECK(hipGetLastError());
    R.odd_frame = false;
}

// ReSTIR_DI_InitialCandidates: output = the initial-candidates buffer.
void restir_initial_pass(Ctx& ctx, uint32_t sample_number, uint32_t random_number, hipStream_t s)
{
    check_ready(ctx);
    RenderState& R = ctx.render;
    const mipt_restir_di_settings& st = R.restir;
    // with an envmap, candidates exist without emissive triangles
    // (InitialCandidates.h:451-459)
    if (ctx.emissive_count > 0 || env_sampling_on(ctx))
    {
        PtData d = make_pt_data(ctx, sample_number, random_number);
        RsData r = make_rs_data(ctx);
        r.out = R.reservoirs[0].as<float4>();
        if (st.do_light_presampling)
        {
            // the records hold triangle indices: never read them unfilled
            if (!R.presampled_ready)
                fail(MIPT_ERR_INVALID_ARGUMENT, "light presampling is on: launch "
                                                "MIPT_PASS_RESTIR_DI_LIGHTS_PRESAMPLING before the initial candidates");
            r.presampled = R.presampled.as<float4>();
        }
        const unsigned g = blocks_for(r.n, 256);
        hipLaunchKernelGGL(k_ic_start, dim3(g), dim3(256), 0, s, d, r);
        for (int32_t j = 0; j < st.number_of_initial_bsdf_candidates; j++)
        {
            trace_closest(ctx, reinterpret_cast<const mipt_ray*>(r.rays), R.rs_hits.as<mipt_hit>(), nullptr, r.n, s);
            hipLaunchKernelGGL(k_ic_bsdf, dim3(g), dim3(256), 0, s, d, r);
        }
        trace_any(ctx, reinterpret_cast<const mipt_ray*>(r.rays + 2 * (size_t)r.n), R.rs_occ.as<uint8_t>() + r.n, r.n,
                  s);
        hipLaunchKernelGGL(k_ic_finish, dim3(g), dim3(256), 0, s, r);
        HIP_CHECK(hipGetLastError());
    }
    if (!st.do_temporal_reuse_pass && !st.do_spatial_reuse_pass && !st.do_fused_spatiotemporal)
        end_frame(R, 0);
}

// ReSTIR_DI_LightsPresampling (ReSTIRDIRenderPass.cpp:411-447): subsets x
// subset_size presampled lights, one thread each.
void restir_presampling_pass(Ctx& ctx, uint32_t sample_number, uint32_t random_number, hipStream_t s)
{
    check_ready(ctx);
    RenderState& R = ctx.render;
    const mipt_restir_di_settings& st = R.restir;
    if (!st.do_light_presampling)
        fail(MIPT_ERR_INVALID_ARGUMENT, "light presampling is off in the ReSTIR DI settings");
    const bool env = env_sampling_on(ctx);
    if (ctx.emissive_count == 0 && !env)
        return; // LightsPresampling.h:80-82: nothing to presample
    const float env_prob = env ? (ctx.emissive_count == 0 ? 1.0f : st.envmap_candidate_probability) : 0.0f;
    const uint32_t count = (uint32_t)st.light_presampling_number_of_subsets * (uint32_t)st.light_presampling_subset_size;
    PtData d = make_pt_data(ctx, sample_number, random_number);
    hipLaunchKernelGGL(k_presample_lights, dim3(blocks_for(count, 256)), dim3(256), 0, s, d, env_prob,
                       R.presampled.as<float4>(), count);
    HIP_CHECK(hipGetLastError());
    R.presampled_ready = true;
}

// ReSTIR_DI_SpatiotemporalReuse (ReSTIRDIRenderPass.cpp:624-657 for the
// buffers): reads last frame's output, writes the other spatial buffer.
void restir_spatiotemporal_pass(Ctx& ctx, uint32_t sample_number, uint32_t random_number, hipStream_t s)
{
    check_ready(ctx);
    RenderState& R = ctx.render;
    const mipt_restir_di_settings& st = R.restir;
    if (!st.do_fused_spatiotemporal)
        fail(MIPT_ERR_INVALID_ARGUMENT, "the fused spatiotemporal pass is off in the ReSTIR DI settings");
    const int32_t in = R.last_output;
    const int32_t out = in == 1 ? 2 : 1;
    PtData d = make_pt_data(ctx, sample_number, random_number);
    RsData r = make_rs_data(ctx);
    r.in = R.reservoirs[in].as<float4>();
    r.init = R.reservoirs[0].as<float4>();
    r.out = R.reservoirs[out].as<float4>();
    r.pass_index = 0;
    r.kill_after = st.do_temporal_reuse_pass || st.spatial_number_of_passes - 1 != 0;
    const int32_t K = st.spatial_reuse_neighbor_count;
    const unsigned g = blocks_for(r.n, 256);
    hipLaunchKernelGGL(k_fst_start, dim3(g), dim3(256), 0, s, d, r);
    trace_any(ctx, reinterpret_cast<const mipt_ray*>(r.rays), R.rs_occ.as<uint8_t>(), (uint32_t)(K + 2) * r.n, s);
    hipLaunchKernelGGL(k_fst_combine, dim3(g), dim3(256), 0, s, d, r);
    trace_any(ctx, reinterpret_cast<const mipt_ray*>(r.rays), R.rs_occ.as<uint8_t>(), r.n, s);
    hipLaunchKernelGGL(k_sp_finish, dim3(g), dim3(256), 0, s, r);
    HIP_CHECK(hipGetLastError());
    R.fused_output = out;
    if (st.spatial_number_of_passes == 1)
        end_frame(R, out);
}

// ReSTIR_DI_TemporalReuse (ReSTIRDIRenderPass.cpp:466-503 for the buffers).
void restir_temporal_pass(Ctx& ctx, uint32_t sample_number, uint32_t random_number, hipStream_t s)
{
    check_ready(ctx);
    RenderState& R = ctx.render;
    const mipt_restir_di_settings& st = R.restir;
    if (!st.do_temporal_reuse_pass)
        fail(MIPT_ERR_INVALID_ARGUMENT, "temporal reuse is disabled in the ReSTIR DI settings");
    const int32_t out = st.do_spatial_reuse_pass ? 0 : (R.odd_frame ? 1 : 2);
    PtData d = make_pt_data(ctx, sample_number, random_number);
    RsData r = make_rs_data(ctx);
    r.in = R.reservoirs[R.last_output].as<float4>();
    r.init = R.reservoirs[0].as<float4>();
    r.out = R.reservoirs[out].as<float4>();
    const unsigned g = blocks_for(r.n, 256);
    hipLaunchKernelGGL(k_tp_start, dim3(g), dim3(256), 0, s, d, r);
    trace_any(ctx, reinterpret_cast<const mipt_ray*>(r.rays), R.rs_occ.as<uint8_t>(), 2 * r.n, s);
    hipLaunchKernelGGL(k_tp_finish, dim3(g), dim3(256), 0, s, d, r);
    HIP_CHECK(hipGetLastError());
    R.temporal_output = out;
    if (!st.do_spatial_reuse_pass)
        end_frame(R, out);
}

// ReSTIR_DI_SpatialReuse (ReSTIRDIRenderPass.cpp:530-571 for the buffers).
void restir_spatial_pass(Ctx& ctx, int32_t pass_index, uint32_t sample_number, uint32_t random_number,
                         hipStream_t s)
{
    check_ready(ctx);
    RenderState& R = ctx.render;
    const mipt_restir_di_settings& st = R.restir;
    // the fused pass is followed by spatial passes 1 .. P-1 whatever
    // do_spatial_reuse_pass says (ReSTIRDIRenderPass.cpp:634-656)
    if (!st.do_spatial_reuse_pass && !st.do_fused_spatiotemporal)
        fail(MIPT_ERR_INVALID_ARGUMENT, "spatial reuse is disabled in the ReSTIR DI settings");
    if (pass_index < 0 || pass_index >= st.spatial_number_of_passes)
        fail(MIPT_ERR_INVALID_ARGUMENT, "spatial pass index out of range");
    int32_t in, out;
    if (st.do_fused_spatiotemporal)
    {
        // configure_spatial_pass_for_fused_spatiotemporal
        // (ReSTIRDIRenderPass.cpp:573-604): pass 0 is the fused pass; later
        // passes read the previous one's output and write the other buffer
        if (pass_index == 0)
            fail(MIPT_ERR_INVALID_ARGUMENT, "with do_fused_spatiotemporal, spatial pass 0 is the fused pass");
        in = R.fused_output;
        out = in == 1 ? 2 : 1;
    }
    else if (pass_index == 0)
    {
        in = st.do_temporal_reuse_pass ? R.temporal_output : 0;
        out = 1;
    }
    else if ((pass_index & 1) == 0)
    {
        in = 2;
        out = 1;
    }
    else
    {
        in = 1;
        out = 2;
    }
    const bool last = pass_index == st.spatial_number_of_passes - 1;
    PtData d = make_pt_data(ctx, sample_number, random_number);
    RsData r = make_rs_data(ctx);
    r.in = R.reservoirs[in].as<float4>();
    r.out = R.reservoirs[out].as<float4>();
    r.pass_index = pass_index;
    // SpatialReuse.h:304-314: visibility reuse unless this is the final pass
    // of a frame without temporal reuse.
    r.kill_after = st.do_temporal_reuse_pass || !last;
    const unsigned g = blocks_for(r.n, 256);
    const int32_t K = st.spatial_reuse_neighbor_count;
    hipLaunchKernelGGL(k_sp_start, dim3(g), dim3(256), 0, s, d, r);
    if (K > 0)
        trace_any(ctx, reinterpret_cast<const mipt_ray*>(r.rays), R.rs_occ.as<uint8_t>(), (uint32_t)K * r.n, s);
    hipLaunchKernelGGL(k_sp_center, dim3(g), dim3(256), 0, s, d, r);
    trace_any(ctx, reinterpret_cast<const mipt_ray*>(r.rays), R.rs_occ.as<uint8_t>(), r.n, s);
    hipLaunchKernelGGL(k_sp_finish, dim3(g), dim3(256), 0, s, r);
    HIP_CHECK(hipGetLastError());
    if (st.do_fused_spatiotemporal)
        R.fused_output = out;
    if (last)
        end_frame(R, out);
}

} // namespace mipt
