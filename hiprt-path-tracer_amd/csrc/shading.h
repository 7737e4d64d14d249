// shading.h -- materials, the layered Principled BSDF (diffuse + metal lobes),
// light sampling, Russian roulette and G-buffer normal packing, as device
// functions shared by the path-tracing kernels.
//
// Every function restates the reference's arithmetic operation for operation
// (compiled with -ffp-contract=off). sin, cos and pow go through the
// fixed-sequence versions of pt_math.h, so a CPU restatement of the same
// steps reproduces every result bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pt_math.h"
#include "volume.h"

namespace pt {

constexpr float PI_F = 3.14159265358979323846f;
constexpr float TWO_PI_F = 6.28318530717958647693f;
constexpr float INV_PI_F = 0.31830988618379067154f;
constexpr float ROUGHNESS_CLAMP = 1.0e-4f;                 // MaterialConstants.h:21
constexpr float PERFECTLY_SMOOTH_THRESHOLD = 1.0e-2f;      // MaterialConstants.h:22
constexpr float DELTA_HIGH_VALUE = 1.0e9f;                 // MaterialConstants.h:23
constexpr float DELTA_ALIGN_THRESHOLD = 0.999999f;         // MaterialConstants.h:24
constexpr float GGX_DOT_CLAMP = 1.0e-3f;                   // Microfacet.h:20

// hippt::max / min / clamp (src/HostDeviceCommon/Math.h:373-407).
PT_HD float maxf_(float a, float b) { return a > b ? a : b; }
PT_HD float minf_(float a, float b) { return a < b ? a : b; }
PT_HD float clampf_(float lo, float hi, float v) { return minf_(hi, maxf_(lo, v)); }
PT_HD float lerpf_(float a, float b, float t) { return (1.0f - t) * a + t * b; }

struct Color {
    float r, g, b;
};
PT_HD Color col(float r, float g, float b) { return Color{r, g, b}; }
PT_HD Color col(float v) { return Color{v, v, v}; }
PT_HD Color operator+(Color a, Color b) { return col(a.r + b.r, a.g + b.g, a.b + b.b); }
PT_HD Color operator-(Color a, Color b) { return col(a.r - b.r, a.g - b.g, a.b - b.b); }
PT_HD Color operator*(Color a, Color b) { return col(a.r * b.r, a.g * b.g, a.b * b.b); }
PT_HD Color operator*(Color c, float k) { return col(c.r * k, c.g * k, c.b * k); }
PT_HD Color operator*(float k, Color c) { return col(c.r * k, c.g * k, c.b * k); }
PT_HD Color operator/(Color c, float k) { return col(c.r / k, c.g / k, c.b / k); }
PT_HD Color operator/(Color a, Color b) { return col(a.r / b.r, a.g / b.g, a.b / b.b); }
PT_HD float luminance(Color c) { return 0.3086f * c.r + 0.6094f * c.g + 0.0820f * c.b; } // Color.h:90
PT_HD float max_component(Color c) { return maxf_(c.r, maxf_(c.g, c.b)); }
PT_HD bool is_black(Color c) { return !(c.r > 0.0f || c.g > 0.0f || c.b > 0.0f); }
PT_HD Color lerpc(Color a, Color b, float t) { return (1.0f - t) * a + t * b; }

// Unpacked effective material (after the reference's 8-bit packing round
// trip, MaterialPacked.h, and the coat roughening of get_intersection_material,
// Material.h:122-149). 256 bytes.
struct Material {
    Color emission;
    Color base_color;
    float roughness;
    float metallic;
    Color metallic_F82;
    Color metallic_F90;
    float metallic_F90_falloff_exponent;
    float anisotropy;
    float anisotropy_rotation;
    float second_roughness_weight;
    float second_roughness;
    float specular;
    float coat;
    float sheen;
    float specular_transmission;
    float diffuse_transmission;
    float thin_film;
    float ior;
    uint32_t flags; // MAT_* bits
    float specular_tint;
    Color specular_color;
    float specular_darkening;
    Color coat_medium_absorption;
    float coat_medium_thickness;
    float coat_roughness;
    float coat_roughening;
    float coat_darkening;
    float coat_anisotropy;
    float coat_anisotropy_rotation;
    float coat_ior;
    float sheen_roughness;
    Color sheen_color;
    float absorption_at_distance;
    Color absorption_color;
    float dispersion_scale;
    float dispersion_abbe_number;
    int32_t dielectric_priority;
    float thin_film_ior;
    float thin_film_thickness;       // nm
    float thin_film_kappa_3;         // extinction of the base under the film (IOR override only)
    float thin_film_hue_shift;       // hue shift / 360 degrees, 8-bit
    float thin_film_base_ior_override;
    float pad[6];
};
static_assert(sizeof(Material) == 256, "Material must be 256 bytes");

constexpr uint32_t MAT_METALLIC_EC = 1u;
constexpr uint32_t MAT_SPECULAR_EC = 2u;
constexpr uint32_t MAT_COAT_EC = 4u;
constexpr uint32_t MAT_GLASS_EC = 8u;
constexpr uint32_t MAT_THIN_WALLED = 16u;
constexpr uint32_t MAT_THIN_FILM_IOR_OVERRIDE = 32u;

// Sampled-lobe information (src/Device/includes/BSDFs/BSDFIncidentLightInfo.h).
enum LobeInfo : uint32_t {
    NO_INFO = 0,
    FROM_COAT = 1u << 1,
    FROM_FIRST_METAL = 1u << 2,
    FROM_SECOND_METAL = 1u << 3,
    FROM_SPECULAR = 1u << 4,
    FROM_GLASS_REFLECT = 1u << 5,
    FROM_GLASS_REFRACT = 1u << 6,
    FROM_DIFFUSE = 1u << 7,
    FROM_DIFFUSE_TRANSMISSION = 1u << 8,
};

// ReSTIRDISampleFlags (src/Device/includes/ReSTIR/DI/SampleFlags.h): an 8-bit
// field holding the sampled lobe's LobeInfo bits and, in bit 7, UNOCCLUDED
// (which FROM_DIFFUSE shares; FROM_DIFFUSE_TRANSMISSION does not fit).
// flags_to_BSDF_incident_light_info keeps bits 1-6 (DI/Reservoir.h:45-48).
constexpr uint32_t RESTIR_FLAG_UNOCCLUDED = 1u << 7;
// RESTIR_DI_FLAGS_ENVMAP_SAMPLE (ReSTIR/DI/SampleFlags.h:16): the sample's
// point is a direction in envmap space
constexpr uint32_t RESTIR_FLAG_ENVMAP = 1u << 0;
constexpr uint32_t RESTIR_FLAG_LOBE_MASK = 0x3fu << 1;

enum RegMode : uint32_t { NO_REGULARIZATION = 0, REGULARIZATION_CLASSIC = 1, REGULARIZATION_MIS = 2 };

// Per-pass constants of the BSDF (BSDFsData.h + MicrofacetRegularizationSettings.h).
struct BsdfGlobals {
    const float* ggx_conductor_ess; // 128 x 128, row 0 = roughness 1
    const float* glossy_dielectric_ess; // [ior 128][roughness 64][cos 128], rows flipped like the 2D LUT
    const float* sheen_ltc;         // 32 x 32 float3 (Ai, Bi, Ri)
    const float* glass_ess;         // [F0^(1/4) 128][roughness 16][cos 256], entering the medium
    const float* glass_inv_ess;     // same, leaving the medium
    const float* thin_glass_ess;    // [F0^(1/4) 96][roughness 32][cos 32]
    const float* glass_exponent;    // [lower/higher][IOR knot 10][roughness segment 11][r_max, kind, a, b]
    const Material* mats;           // media of the nested-dielectrics stack (IOR, absorption)
    int32_t lut_size;
    float consistent_tau;           // MicrofacetRegularization::consistent_tau(tau_0, sample_number)
    float consistent_tau_refract;   // consistent_tau(tau_0, sample_number + 1)
    float min_roughness;
    int32_t regularize;             // PrincipledBSDFDoMicrofacetRegularization
};

PT_HD f3 vmul(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
PT_HD f3 vdiv(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }

// ONB.h:18-30
PT_HD void build_onb(f3 N, f3& T, f3& B)
{
    if (N.z < -0.99998796f)
    {
        T = mk(0.0f, -1.0f, 0.0f);
        B = mk(-1.0f, 0.0f, 0.0f);
        return;
    }
    float nxa = -N.x / (1.0f + N.z);
    T = mk(1.0f + N.x * nxa, nxa * N.y, -N.x);
    B = mk(T.y, 1.0f - N.y * N.y / (1.0f + N.z), -N.y);
}

// ONB.h:35-43; rotation angle is anisotropy_rotation * PI.
PT_HD void build_rotated_onb(f3 N, f3& T, f3& B, float rotation)
{
    f3 up = fabsf(N.z) < 0.9999999f ? mk(0.0f, 0.0f, 1.0f) : mk(1.0f, 0.0f, 0.0f);
    T = normalize(cross(up, N));
    float c = cos_fixed(rotation), s = sin_fixed(rotation);
    T = (vmul(T, c) + vmul(cross(N, T), s)) + vmul(vmul(N, dot(N, T)), 1.0f - c);
    B = cross(N, T);
}

PT_HD f3 world_to_local(f3 T, f3 B, f3 N, f3 V) { return normalize(mk(dot(V, T), dot(V, B), dot(V, N))); }
PT_HD f3 local_to_world(f3 T, f3 B, f3 N, f3 V) { return normalize((vmul(T, V.x) + vmul(B, V.y)) + vmul(N, V.z)); }

PT_HD f3 reflect_ray(f3 d, f3 n)
{
    float k = 2.0f * dot(d, n);
    return vmul(n, k) - d;
}

// Sampling.h:180-201
PT_HD f3 cosine_sample_around_normal(f3 n, Xorshift32& rng)
{
    float r1 = rng();
    float r2 = 2.0f * rng() - 1.0f;
    if (r1 < 1.0e-8f && r2 < -0.999999f && n.z > 0.999999f)
    {
        r1 += 1.0e-7f;
        r2 += 1.0e-7f;
    }
    float theta = TWO_PI_F * r1;
    float s = sqrtf(1.0f - r2 * r2);
    f3 sp = mk(s * cos_fixed(theta), s * sin_fixed(theta), r2);
    return normalize(n + sp);
}

// MaterialUtils.h:25-30
PT_HD void get_alphas(float roughness, float anisotropy, float& ax, float& ay)
{
    float aspect = sqrtf(1.0f - 0.9f * anisotropy);
    ax = maxf_(ROUGHNESS_CLAMP, roughness * roughness / aspect);
    ay = maxf_(ROUGHNESS_CLAMP, roughness * roughness * aspect);
}

PT_HD bool is_perfectly_smooth(float r) { return r <= PERFECTLY_SMOOTH_THRESHOLD; }

enum DeltaSampled : int { NOT_SPECULAR = 0, SPECULAR_PEAK_SAMPLED = 1, SPECULAR_PEAK_NOT_SAMPLED = 2 };

// MaterialUtils.h:143-175: whether a perfectly smooth lobe can evaluate to
// anything but 0 for the incident direction `info` describes.
PT_HD DeltaSampled delta_reflection_sampled(const Material& m, float roughness, float anisotropy, uint32_t info)
{
    if (!is_perfectly_smooth(roughness))
        return NOT_SPECULAR;
    bool matching = fabsf(anisotropy - m.anisotropy) < 1.0e-3f;
    bool glass = info == FROM_GLASS_REFLECT && is_perfectly_smooth(m.roughness) && matching;
    if (glass)
        return SPECULAR_PEAK_SAMPLED;
    bool first = info == FROM_FIRST_METAL && is_perfectly_smooth(m.roughness) && matching;
    bool second = info == FROM_SECOND_METAL && is_perfectly_smooth(m.second_roughness) && matching;
    if (first || second)
        return SPECULAR_PEAK_SAMPLED;
    bool coat_matching = fabsf(anisotropy - m.coat_anisotropy) < 1.0e-3f;
    bool coat = info == FROM_COAT && coat_matching && is_perfectly_smooth(m.coat_roughness);
    if (coat)
        return SPECULAR_PEAK_SAMPLED;
    bool specular = info == FROM_SPECULAR && is_perfectly_smooth(m.roughness) && matching;
    if (specular)
        return SPECULAR_PEAK_SAMPLED;
    return SPECULAR_PEAK_NOT_SAMPLED;
}

// MicrofacetRegularization.h:22-45
PT_HD float regularize_reflection(const BsdfGlobals& g, uint32_t mode, float roughness, float path_roughness)
{
    if (!g.regularize || mode == NO_REGULARIZATION)
        return roughness;
    float path_tau = g.consistent_tau / maxf_(path_roughness * path_roughness, 1.0e-8f);
    float reg = sqrtf(sqrtf(1.0f / (path_tau * PI_F)));
    return maxf_(g.min_roughness, maxf_(roughness, reg));
}

// Microfacet.h:25-69
PT_HD float ggx_d(float ax, float ay, f3 h)
{
    float denom = (h.x * h.x) / (ax * ax) + (h.y * h.y) / (ay * ay) + (h.z * h.z);
    return 1.0f / (PI_F * ax * ay * denom * denom);
}
PT_HD float ggx_vndf(float D, float G1V, f3 v, f3 h)
{
    float HoL = maxf_(GGX_DOT_CLAMP, dot(v, h));
    return G1V * D * HoL / v.z;
}
PT_HD float smith_lambda(float ax, float ay, f3 d)
{
    float a = d.x * ax, b = d.y * ay;
    return (-1.0f + sqrtf(1.0f + (a * a + b * b) / (d.z * d.z))) * 0.5f;
}

// LUT lookup of the GGX conductor single-scattering albedo E_ss(cos, roughness),
// nearest texel: column (int)(u * (N-1)), row (int)(v * (N-1)) with
// u = cos, v = 1 - roughness (MicrofacetEnergyCompensation.h:27-45).
PT_HD float lut_fetch(const float* lut, int n, float u, float v)
{
    if (u != 1.0f) u -= (float)(int)u;
    if (v != 1.0f) v -= (float)(int)v;
    u = u < 0.0f ? 1.0f + u : u;
    v = v < 0.0f ? 1.0f + v : v;
    int x = (int)(u * (float)(n - 1));
    int y = (int)(v * (float)(n - 1));
    return lut[x + y * n];
}

// MicrofacetEnergyCompensation.h:27-45; thresholds of BSDFsData.h (0, -1).
PT_HD Color ggx_energy_compensation_conductor(const BsdfGlobals& g, Color F0, float roughness, bool do_ec, f3 lv)
{
    bool smooth_enough = roughness <= 0.0f;
    bool invalid = lv.z < 0.0f;
    if (!do_ec || smooth_enough || invalid)
        return col(1.0f);
    float u = maxf_(0.0f, lv.z);
    float v = 1.0f - roughness;
    float ess = lut_fetch(g.ggx_conductor_ess, g.lut_size, u, v);
    float kms = (1.0f - ess) / ess;
    return col(1.0f) + kms * F0;
}

// Microfacet.h:103-178 (single scattering, height-uncorrelated Smith G2).
PT_HD Color ggx_eval_reflect_ss(float roughness, float anisotropy, Color F, f3 lv, f3 ll, f3 lh, float& pdf,
                                DeltaSampled delta)
{
    pdf = 0.0f;
    if (is_perfectly_smooth(roughness))
    {
        if (delta == SPECULAR_PEAK_NOT_SAMPLED)
            return col(0.0f);
        if (dot(reflect_ray(lv, mk(0.0f, 0.0f, 1.0f)), ll) < DELTA_ALIGN_THRESHOLD)
        {
            pdf = 0.0f;
            return col(0.0f);
        }
        pdf = DELTA_HIGH_VALUE;
        return col(DELTA_HIGH_VALUE) * F / fabsf(ll.z);
    }
    if (ll.z < 0.0f)
        return col(0.0f);
    float ax, ay;
    get_alphas(roughness, anisotropy, ax, ay);
    float D = ggx_d(ax, ay, lh);
    float lambda_v = smith_lambda(ax, ay, lv);
    float G1V = 1.0f / (1.0f + lambda_v);
    float Dvis = ggx_vndf(D, G1V, lv, lh);
    float NoV = maxf_(GGX_DOT_CLAMP, fabsf(lv.z));
    float NoL = maxf_(GGX_DOT_CLAMP, fabsf(ll.z));
    pdf = Dvis / (4.0f * dot(lv, lh));
    if (pdf == 0.0f)
        return col(0.0f);
    float lambda_l = smith_lambda(ax, ay, ll);
    float G1L = 1.0f / (1.0f + lambda_l);
    float G2 = G1V * G1L;
    return F * D * G2 / (4.0f * NoL * NoV);
}

// Microfacet.h:227-287
PT_HD float ggx_pdf_reflect(float roughness, float anisotropy, f3 lv, f3 ll, f3 lh, DeltaSampled delta)
{
    if (is_perfectly_smooth(roughness))
    {
        if (delta == SPECULAR_PEAK_NOT_SAMPLED)
            return 0.0f;
        if (dot(reflect_ray(lv, mk(0.0f, 0.0f, 1.0f)), ll) < DELTA_ALIGN_THRESHOLD)
            return 0.0f;
        return DELTA_HIGH_VALUE;
    }
    if (ll.z < 0.0f)
        return 0.0f;
    float ax, ay;
    get_alphas(roughness, anisotropy, ax, ay);
    float D = ggx_d(ax, ay, lh);
    float lambda_v = smith_lambda(ax, ay, lv);
    float G1V = 1.0f / (1.0f + lambda_v);
    float Dvis = ggx_vndf(D, G1V, lv, lh);
    return Dvis / (4.0f * dot(lv, lh));
}

// Heitz 2018 VNDF sampling (Microfacet.h:398-424).
PT_HD f3 ggx_vndf_sample(f3 v, float ax, float ay, Xorshift32& rng)
{
    float r1 = rng();
    float r2 = rng();
    f3 Vh = normalize(mk(ax * v.x, ay * v.y, v.z));
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    f3 T1 = lensq > 0.0f ? vdiv(mk(-Vh.y, Vh.x, 0.0f), sqrtf(lensq)) : mk(1.0f, 0.0f, 0.0f);
    f3 T2 = cross(Vh, T1);
    float r = sqrtf(r1);
    float phi = TWO_PI_F * r2;
    float t1 = r * cos_fixed(phi);
    float t2 = r * sin_fixed(phi);
    float s = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - s) * sqrtf(1.0f - t1 * t1) + s * t2;
    f3 Nh = (vmul(T1, t1) + vmul(T2, t2)) + vmul(Vh, sqrtf(maxf_(0.0f, 1.0f - t1 * t1 - t2 * t2)));
    return normalize(mk(ax * Nh.x, ay * Nh.y, maxf_(0.0f, Nh.z)));
}

// Microfacet.h:484-500
PT_HD f3 ggx_sample_reflection(float roughness, float anisotropy, f3 lv, Xorshift32& rng)
{
    float below = lv.z < 0.0f ? -1.0f : 1.0f;
    float ax, ay;
    get_alphas(roughness, anisotropy, ax, ay);
    f3 m;
    if (ax <= ROUGHNESS_CLAMP && ay <= ROUGHNESS_CLAMP)
        m = mk(0.0f, 0.0f, 1.0f);
    else
        m = ggx_vndf_sample(vmul(lv, below), ax, ay, rng);
    return normalize(reflect_ray(lv, vmul(m, below)));
}

PT_HD float pow6(float x)
{
    float x2 = x * x;
    float x4 = x2 * x2;
    return x4 * x2;
}

// Fresnel.h:138-168 (Adobe F82-tint).
PT_HD Color f82_tint_fresnel(Color F0, Color F82, Color F90, float falloff, float cos_theta)
{
    Color base = F0 + (F90 - F0) * pow_fixed(1.0f - cos_theta, falloff);
    if (max_component(base) < 1.0e-8f)
        return col(0.0f);
    float lazanyi = cos_theta * pow6(1.0f - cos_theta);
    const float cos_theta_max = 1.0f / 7.0f;
    const float denom_a = cos_theta_max * pow6(1.0f - cos_theta_max);
    Color nume_a = (F0 + (F90 - F0) * pow_fixed(1.0f - cos_theta_max, falloff)) * (col(1.0f) - F82);
    Color a = nume_a / denom_a;
    Color F = base - a * lazanyi;
    return col(clampf_(0.0f, 1.0f, F.r), clampf_(0.0f, 1.0f, F.g), clampf_(0.0f, 1.0f, F.b));
}

struct BsdfCtx {
    f3 view;          // towards the viewer, world space
    f3 sn;            // shading normal
    f3 gn;            // geometric normal
    f3 to_light;      // world space (eval)
    uint32_t info;    // LobeInfo of to_light
    uint32_t reg;     // RegMode
    float path_roughness;
    VolumeState* vs;  // volume state of the path at this vertex (after the push of its material)
    int32_t update;   // sampling the continuation: pops the stack, resets the distance
};

// Principled.h:125-153
PT_HD Color thin_film_fresnel(const Material& m, float ambient_ior, float HoL);

PT_HD Color metallic_eval(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, float roughness, float anisotropy,
                          float incident_ior, f3 lv, f3 ll, f3 lh, float& pdf)
{
    float rr = regularize_reflection(g, c.reg, roughness, c.path_roughness);
    DeltaSampled ds = delta_reflection_sampled(m, rr, anisotropy, c.info);
    if (ds == SPECULAR_PEAK_NOT_SAMPLED)
    {
        pdf = 0.0f;
        return col(0.0f);
    }
    float HoL = clampf_(1.0e-8f, 1.0f, dot(lh, ll));
    Color F_metal = f82_tint_fresnel(m.base_color, m.metallic_F82, m.metallic_F90, m.metallic_F90_falloff_exponent, HoL);
    Color F = lerpc(F_metal, thin_film_fresnel(m, incident_ior, HoL), m.thin_film);
    Color ms = ggx_energy_compensation_conductor(g, F, rr, (m.flags & MAT_METALLIC_EC) != 0, lv);
    Color ss = ggx_eval_reflect_ss(rr, anisotropy, F, lv, ll, lh, pdf, ds);
    return ss * ms;
}

PT_HD float metallic_pdf(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, float roughness, float anisotropy,
                         f3 lv, f3 ll, f3 lh)
{
    float rr = regularize_reflection(g, c.reg, roughness, c.path_roughness);
    DeltaSampled ds = delta_reflection_sampled(m, rr, anisotropy, c.info);
    if (ds == SPECULAR_PEAK_NOT_SAMPLED)
        return 0.0f;
    return ggx_pdf_reflect(rr, anisotropy, lv, ll, lh, ds);
}

// ---------------------------------------------------------------------------
// Dielectric Fresnel (src/Device/includes/Fresnel.h)
// ---------------------------------------------------------------------------
// Fresnel.h:11-26
PT_HD float F0_from_eta(float eta_t, float eta_i)
{
    float nume = eta_t - eta_i;
    float denom = eta_t + eta_i;
    return (nume * nume) / (denom * denom);
}
PT_HD float F0_from_eta_t_and_relative_ior(float eta_t, float relative_eta)
{
    return F0_from_eta(eta_t, eta_t / relative_eta);
}

// Fresnel.h:41-62; relative_eta = eta_t / eta_i.
PT_HD float fresnel_dielectric(float cos_theta_i, float relative_eta)
{
    if (fabsf(1.0f - relative_eta) < 1.0e-4f)
        return 0.0f;
    float sin_theta_i2 = 1.0f - cos_theta_i * cos_theta_i;
    float sin_theta_t2 = sin_theta_i2 / (relative_eta * relative_eta);
    if (sin_theta_t2 >= 1.0f)
        return 1.0f;
    float cos_theta_t = sqrtf(1.0f - sin_theta_t2);
    float r_par = (relative_eta * cos_theta_i - cos_theta_t) / (relative_eta * cos_theta_i + cos_theta_t);
    float r_perp = (cos_theta_i - relative_eta * cos_theta_t) / (cos_theta_i + relative_eta * cos_theta_t);
    return (r_par * r_par + r_perp * r_perp) / 2.0f;
}
PT_HD float fresnel_dielectric(float cos_theta_i, float eta_i, float eta_t)
{
    return fresnel_dielectric(cos_theta_i, eta_t / eta_i);
}

// Fresnel.h:169-172 (d'Eon's fit of the hemispherical albedo of a smooth dielectric).
PT_HD float fresnel_hemispherical_albedo_fit(float relative_eta)
{
    return log_fixed((10893.0f * relative_eta - 1438.2f) /
                     (-774.4f * (relative_eta * relative_eta) + 10212.0f * relative_eta + 1.0f));
}


// ---------------------------------------------------------------------------
// Thin-film interference Fresnel (Belcour & Barla 2017, "A Practical Extension
// to Microfacet Theory for the Modeling of Varying Iridescence", and its
// supplemental document), src/Device/includes/BSDFs/ThinFilm.h
// ---------------------------------------------------------------------------
// eval_sensitivity (ThinFilm.h:13-30): the Gaussian fit of the XYZ colour
// matching functions in the Fourier domain, for an optical path difference in um.
PT_HD Color thin_film_sensitivity(float opd, float shift)
{
    const float phase = 2.0f * PI_F * opd * 1.0e-6f;
    const float val[3] = {5.4856e-13f, 4.4201e-13f, 5.2481e-13f};
    const float pos[3] = {1.6810e+06f, 1.7953e+06f, 2.2084e+06f};
    const float var[3] = {4.3278e+09f, 9.3046e+09f, 6.6121e+09f};
    float xyz[3];
    for (int k = 0; k < 3; k++)
        xyz[k] = val[k] * sqrtf(2.0f * PI_F * var[k]) * cos_fixed(pos[k] * phase + shift) *
                 exp_fixed(-1.0f * var[k] * phase * phase);
    xyz[0] += 9.7470e-14f * sqrtf(2.0f * PI_F * 4.5282e+09f) * cos_fixed(2.2399e+06f * phase + shift) *
              exp_fixed(-4.5282e+09f * phase * phase);
    return col(xyz[0] / 1.0685e-7f, xyz[1] / 1.0685e-7f, xyz[2] / 1.0685e-7f);
}

// fresnel_phase (ThinFilm.h:37-62): phase shifts of the parallel and
// perpendicular polarisations at an interface eta1 | eta2 + i kappa2.
PT_HD void thin_film_phase(float cos_theta_i, float eta1, float eta2, float kappa2, float& phi_par, float& phi_perp)
{
    float sin_theta2 = 1.0f - cos_theta_i * cos_theta_i;
    float A = (eta2 * eta2) * (1.0f - kappa2 * kappa2) - (eta1 * eta1) * sin_theta2;
    float t = 2.0f * (eta2 * eta2) * kappa2;
    float B = sqrtf(A * A + t * t);
    float U = (float)sqrt((double)(A + B) * 0.5); // the reference's 0.5 is a double
    float V = sqrtf((B - A) * 0.5f);
    float ec = eta1 * cos_theta_i;
    float phi_perp_y = 2.0f * eta1 * V * cos_theta_i;
    float phi_perp_x = U * U + V * V - ec * ec;
    phi_perp = atan2_fixed(phi_perp_y, phi_perp_x);
    float phi_par_y = 2.0f * eta1 * (eta2 * eta2) * cos_theta_i * (2.0f * kappa2 * U - (1.0f - kappa2 * kappa2) * V);
    float px = (eta2 * eta2) * (1.0f + kappa2 * kappa2) * cos_theta_i;
    float phi_par_x = px * px - (eta1 * eta1) * (U * U + V * V);
    phi_par = atan2_fixed(phi_par_y, phi_par_x);
}

// fresnel_conductor (ThinFilm.h:64-88): reflectances of both polarisations.
PT_HD void thin_film_conductor(float cos_theta_i, float eta, float k, float& Rp2, float& Rs2)
{
    float c2 = cos_theta_i * cos_theta_i;
    float s2 = 1.0f - c2;
    float temp1 = eta * eta - k * k - s2;
    float a2pb2 = sqrtf(temp1 * temp1 + 4.0f * k * k * eta * eta);
    float a = sqrtf(0.5f * (a2pb2 + temp1));
    float term1 = a2pb2 + c2;
    float term2 = 2.0f * a * cos_theta_i;
    Rs2 = clampf_(0.0f, 1.0f, (term1 - term2) / (term1 + term2));
    float term3 = a2pb2 * c2 + s2 * s2;
    float term4 = term2 * s2;
    Rp2 = clampf_(0.0f, 1.0f, Rs2 * (term3 - term4) / (term3 + term4));
}

// RGB_hue_shift (ThinFilm.h:93-124): rotation about the grey axis.
PT_HD Color rgb_hue_shift(Color c, float degrees)
{
    if (degrees == 0.0f)
        return c;
    float cosA = cos_fixed(degrees / 180.0f * PI_F);
    float sinA = sin_fixed(degrees / 180.0f * PI_F);
    const float sqrt_1_3 = 0.57735026918962576451f;
    const float one_over_3 = 1.0f / 3.0f;
    float m00 = cosA + (1.0f - cosA) / 3.0f;
    float m01 = one_over_3 * (1.0f - cosA) - sqrt_1_3 * sinA;
    float m02 = one_over_3 * (1.0f - cosA) + sqrt_1_3 * sinA;
    float m10 = one_over_3 * (1.0f - cosA) + sqrt_1_3 * sinA;
    float m11 = cosA + one_over_3 * (1.0f - cosA);
    float m12 = one_over_3 * (1.0f - cosA) - sqrt_1_3 * sinA;
    float m20 = one_over_3 * (1.0f - cosA) - sqrt_1_3 * sinA;
    float m21 = one_over_3 * (1.0f - cosA) + sqrt_1_3 * sinA;
    float m22 = cosA + one_over_3 * (1.0f - cosA);
    Color h = col(c.r * m00 + c.g * m01 + c.b * m02, c.r * m10 + c.g * m11 + c.b * m12,
                  c.r * m20 + c.g * m21 + c.b * m22);
    return col(clampf_(0.0f, 1.0f, h.r), clampf_(0.0f, 1.0f, h.g), clampf_(0.0f, 1.0f, h.b));
}

// thin_film_fresnel (ThinFilm.h:131-226): an airy summation of the two
// interfaces ambient | film | base, both polarisations, two orders of the
// spectral terms, then XYZ -> sRGB.
PT_HD Color thin_film_fresnel(const Material& m, float ambient_ior, float HoL)
{
    if (m.thin_film == 0.0f)
        return col(0.0f);
    const bool over = (m.flags & MAT_THIN_FILM_IOR_OVERRIDE) != 0;
    float eta1 = ambient_ior;
    float eta2 = m.thin_film_ior;
    float eta3 = over ? m.thin_film_base_ior_override : m.ior;
    float kappa3 = over ? m.thin_film_kappa_3 : 0.0f;
    float R12p = 0.0f, R12s = 0.0f, T121p = 0.0f, T121s = 0.0f, R23p = 0.0f, R23s = 0.0f;
    float cos_theta_2 = 0.0f;
    float e12 = eta1 / eta2;
    float ct2 = 1.0f - (1.0f - HoL * HoL) * (e12 * e12);
    if (ct2 <= 0.0f)
    {
        R12s = 1.0f;
        R12p = 1.0f;
        T121p = 0.0f;
        T121s = 0.0f;
    }
    else
    {
        cos_theta_2 = sqrtf(ct2);
        thin_film_conductor(HoL, eta2 / eta1, 0.0f, R12p, R12s);
        thin_film_conductor(cos_theta_2, eta3 / eta2, kappa3, R23p, R23s);
        T121p = 1.0f - R12p;
        T121s = 1.0f - R12s;
    }
    float D = m.thin_film_thickness / 1000.0f * cos_theta_2;
    float phi21p, phi21s, phi23p, phi23s;
    thin_film_phase(HoL, eta1, eta2, 0.0f, phi21p, phi21s);
    thin_film_phase(cos_theta_2, eta2, eta3, kappa3, phi23p, phi23s);
    phi21p = PI_F - phi21p;
    phi21s = PI_F - phi21s;
    float r123p = sqrtf(R12p * R23p);
    float r123s = sqrtf(R12s * R23s);
    float Rs = (T121p * T121p * R23p) / (1.0f - R12p * R23p);
    Color I = col(R12p + Rs);
    float Cm = Rs - T121p;
    for (int k = 1; k <= 2; k++)
    {
        Cm *= r123p;
        Color Sm = 2.0f * thin_film_sensitivity((float)k * D, (float)k * (phi23p + phi21p));
        I = I + Cm * Sm;
    }
    Rs = (T121s * T121s * R23s) / (1.0f - R12s * R23s);
    I = I + col(R12s + Rs);
    Cm = Rs - T121s;
    for (int k = 1; k <= 2; k++)
    {
        Cm *= r123s;
        Color Sm = 2.0f * thin_film_sensitivity((float)k * D, (float)k * (phi23s + phi21s));
        I = I + Cm * Sm;
    }
    I = I * 0.5f;
    float r = 2.3646381f * I.r - 0.8965361f * I.g - 0.4680737f * I.b;
    float gg = -0.5151664f * I.r + 1.4264000f * I.g + 0.0887608f * I.b;
    float b = 0.0052037f * I.r - 0.0144081f * I.g + 1.0092106f * I.b;
    I = col(clampf_(0.0f, 1.0f, r), clampf_(0.0f, 1.0f, gg), clampf_(0.0f, 1.0f, b));
    return rgb_hue_shift(I, m.thin_film_hue_shift * 360.0f);
}

PT_HD Color cdiv(float a, Color b) { return col(a / b.r, a / b.g, a / b.b); }
PT_HD bool is_white(Color c) { return c.r == 1.0f && c.g == 1.0f && c.b == 1.0f; }

// Nearest-texel fetch of a stacked 3D LUT as the reference's CPU sampler does it
// (Image32Bit3D::sample_rgba32f, src/Image/Image.cpp:761-803): wrap to [0, 1]
// (1 stays 1), v flipped, texel (int)(u (W-1)), (int)(v (H-1)), layer (int)(w (D-1)).
PT_HD float lut3d_fetch(const float* lut, int w, int h, int d, float u, float v, float t)
{
    if (u != 1.0f) u -= (float)(int)u;
    if (v != 1.0f) v -= (float)(int)v;
    if (t != 1.0f) t -= (float)(int)t;
    u = u < 0.0f ? 1.0f + u : u;
    v = v < 0.0f ? 1.0f + v : v;
    t = t < 0.0f ? 1.0f + t : t;
    v = 1.0f - v;
    int x = (int)(u * (float)(w - 1));
    int y = (int)(v * (float)(h - 1));
    int z = (int)(t * (float)(d - 1));
    return lut[(size_t)z * (size_t)(w * h) + (size_t)(x + y * w)];
}

// ---------------------------------------------------------------------------
// Specular (dielectric) layer over the diffuse base: the "glossy base"
// ---------------------------------------------------------------------------
// Principled.h:249-285: IOR of the specular layer over the layer above it.
PT_HD float specular_relative_ior(const Material& m, float incident_medium_ior)
{
    if (m.coat == 0.0f)
        return m.ior;
    float incident_layer_ior = lerpf_(incident_medium_ior, m.coat_ior, m.coat);
    float relative_ior = m.ior / incident_layer_ior;
    if (relative_ior < 1.0f)
        relative_ior = 1.0f / relative_ior;
    return relative_ior;
}

// Principled.h:215-241: the dielectric Fresnel lerped with the thin-film one
// (seen from the layer above the specular layer).
PT_HD Color specular_fresnel(const Material& m, float relative_specular_ior, float cos_theta_i)
{
    float layer_above_ior = m.ior / relative_specular_ior;
    Color F_specular = col(0.0f);
    if (m.thin_film < 1.0f)
        F_specular = col(fresnel_dielectric(cos_theta_i, relative_specular_ior));
    Color F_thin_film = thin_film_fresnel(m, layer_above_ior, cos_theta_i);
    return lerpc(F_specular, F_thin_film, m.thin_film);
}

// PrincipledEnergyCompensation.h:15-60 (roughness threshold 0, no bounce limit).
PT_HD float energy_compensation_glossy_base(const BsdfGlobals& g, const Material& m, float incident_medium_ior,
                                            float NoV)
{
    bool disabled = !(m.flags & MAT_SPECULAR_EC);
    bool roughness_low_enough = m.roughness < 0.0f;
    bool no_specular_layer = m.specular < 1.0e-3f;
    bool invalid_view_direction = NoV < 0.0f;
    if (disabled || roughness_low_enough || no_specular_layer || invalid_view_direction)
        return 1.0f;
    float ior = m.ior;
    float relative_ior = specular_relative_ior(m, incident_medium_ior);
    if (fabsf(relative_ior - 1.0f) < 1.0e-3f)
        relative_ior += 1.0e-3f;
    float view_dir_remapped = pow_fixed(NoV, 1.0f / 2.5f);
    float F0_remapped = sqrtf(sqrtf(F0_from_eta_t_and_relative_ior(ior, relative_ior)));
    float ms = lut3d_fetch(g.glossy_dielectric_ess, 128, 64, 128, view_dir_remapped, m.roughness, F0_remapped);
    float compensation = lerpf_(1.0f, ms, m.specular);
    compensation = lerpf_(compensation, 1.0f, m.thin_film);
    return compensation;
}

// PrincipledEnergyCompensation.h:76-125 (the whole BSDF under the coat).
PT_HD float energy_compensation_clearcoat(const BsdfGlobals& g, const Material& m, float incident_medium_ior, float NoV)
{
    bool disabled = !(m.flags & MAT_SPECULAR_EC);
    bool no_coat_layer = m.coat < 1.0e-3f;
    bool invalid_view_direction = NoV < 0.0f;
    if (disabled || no_coat_layer || invalid_view_direction)
        return 1.0f;
    if (fabsf(m.coat_ior / incident_medium_ior - 1.0f) < 1.0e-3f)
        incident_medium_ior += 1.0e-3f;
    float view_dir_remapped = pow_fixed(NoV, 1.0f / 2.5f);
    float F0_remapped = sqrtf(sqrtf(F0_from_eta(m.coat_ior, incident_medium_ior)));
    float ms = lut3d_fetch(g.glossy_dielectric_ess, 128, 64, 128, view_dir_remapped, m.coat_roughness, F0_remapped);
    float compensation = lerpf_(1.0f, ms, m.coat * (1.0f - m.specular_transmission));
    compensation = lerpf_(compensation, 1.0f, m.thin_film);
    return compensation;
}

// Principled.h:292-313
PT_HD Color specular_eval(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, float relative_ior, f3 lv, f3 ll,
                          f3 lh, float& pdf)
{
    float rr = regularize_reflection(g, c.reg, m.roughness, c.path_roughness);
    DeltaSampled ds = delta_reflection_sampled(m, rr, m.anisotropy, c.info);
    Color F = specular_fresnel(m, relative_ior, dot(ll, lh));
    return ggx_eval_reflect_ss(rr, m.anisotropy, F, lv, ll, lh, pdf, ds);
}

PT_HD float specular_pdf(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, f3 lv, f3 ll, f3 lh)
{
    float rr = regularize_reflection(g, c.reg, m.roughness, c.path_roughness);
    DeltaSampled ds = delta_reflection_sampled(m, rr, m.anisotropy, c.info);
    return ggx_pdf_reflect(rr, m.anisotropy, lv, ll, lh, ds);
}

// Principled.h:1272-1303 (OpenPBR's TIR darkening with a diffuse base).
PT_HD Color specular_darkening(const Material& m, float relative_eta, float view_dir_fresnel)
{
    if (m.specular_darkening == 0.0f)
        return col(1.0f);
    float Kr = 1.0f - (1.0f - fresnel_hemispherical_albedo_fit(relative_eta)) / (relative_eta * relative_eta);
    float K = Kr;
    Color base_albedo = m.base_color;
    Color darkening = cdiv(1.0f - K, col(1.0f) - base_albedo * K);
    return lerpc(col(1.0f), darkening, m.specular * m.specular_darkening * (1.0f - m.diffuse_transmission));
}

// ---------------------------------------------------------------------------
// Coat layer
// ---------------------------------------------------------------------------
// Principled.h:49-73
PT_HD Color coat_eval(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, f3 lv, f3 ll, f3 lh,
                      float incident_medium_ior, float& pdf)
{
    float HoL = clampf_(1.0e-8f, 1.0f, dot(lh, ll));
    float rr = regularize_reflection(g, c.reg, m.coat_roughness, c.path_roughness);
    DeltaSampled ds = delta_reflection_sampled(m, rr, m.coat_anisotropy, c.info);
    Color F = col(fresnel_dielectric(HoL, incident_medium_ior, m.coat_ior));
    return ggx_eval_reflect_ss(rr, m.coat_anisotropy, F, lv, ll, lh, pdf, ds);
}

PT_HD float coat_pdf(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, f3 lv, f3 ll, f3 lh)
{
    float rr = regularize_reflection(g, c.reg, m.coat_roughness, c.path_roughness);
    DeltaSampled ds = delta_reflection_sampled(m, rr, m.coat_anisotropy, c.info);
    return ggx_pdf_reflect(rr, m.coat_anisotropy, lv, ll, lh, ds);
}

// Principled.h:890-950 (OpenPBR coat darkening).
PT_HD Color coat_darkening(const Material& m, float relative_eta, float view_dir_fresnel)
{
    if (m.coat_darkening == 0.0f)
        return col(1.0f);
    float Kr = 1.0f - (1.0f - fresnel_hemispherical_albedo_fit(relative_eta)) / (relative_eta * relative_eta);
    float Ks = view_dir_fresnel;
    float base_roughness = lerpf_(1.0f, m.roughness, maxf_(m.specular_transmission, maxf_(m.metallic, m.specular)));
    float K = lerpf_(Ks, Kr, base_roughness);
    float sheen = m.sheen;
    Color base_albedo = (m.base_color + m.sheen_color * sheen) / (1.0f + sheen);
    Color darkening = cdiv(1.0f - K, col(1.0f) - base_albedo * K);
    return lerpc(col(1.0f), darkening, m.coat * m.coat_darkening * (1.0f - m.diffuse_transmission));
}

// ---------------------------------------------------------------------------
// Sheen: LTC fit of the SGGX volumetric sheen (src/Device/includes/BSDFs/SheenLTC.h)
// ---------------------------------------------------------------------------
// read_LTC_parameters through the CPU 2D sampler (Image32Bit::sample_rgba32f,
// src/Image/Image.cpp:564-600): uv = (cos_theta, clamp(roughness)), v flipped.
PT_HD Color sheen_ltc_params(const BsdfGlobals& g, float roughness, float cos_theta)
{
    float u = cos_theta, v = clampf_(0.0f, 1.0f, roughness);
    if (u != 1.0f) u -= (float)(int)u;
    if (v != 1.0f) v -= (float)(int)v;
    u = u < 0.0f ? 1.0f + u : u;
    v = v < 0.0f ? 1.0f + v : v;
    v = 1.0f - v;
    int x = (int)(u * 31.0f);
    int y = (int)(v * 31.0f);
    const float* p = g.sheen_ltc + 3 * (x + y * 32);
    return col(p[0], p[1], p[2]);
}

// SheenLTC.h:24-45
PT_HD float eval_ltc(f3 l, Color AiBiRi)
{
    f3 o = mk(l.x * AiBiRi.r + l.z * AiBiRi.g, l.y * AiBiRi.r, l.z);
    float len = sqrtf(dot(o, o));
    o = vdiv(o, len);
    float det = AiBiRi.r * AiBiRi.r;
    float jacobian = det / (len * len * len);
    return o.z * INV_PI_F * jacobian;
}

// SheenLTC.h:63-70
PT_HD float get_phi(f3 d)
{
    float p = atan2_fixed(d.y, d.x);
    if (p < 0.0f)
        p += TWO_PI_F;
    return p;
}

// SheenLTC.h:75-81: rotation of vec by angle around axis (Rodrigues).
PT_HD f3 rotate_vector(f3 vec, f3 axis, float angle)
{
    float s, co;
    sincos_fixed(angle, s, co);
    return (vmul(vec, co) + vmul(vmul(axis, dot(vec, axis)), 1.0f - co)) + vmul(cross(axis, vec), s);
}

// SheenLTC.h:88-125; lv, ll in the (unrotated) local shading frame.
PT_HD Color sheen_eval(const BsdfGlobals& g, const Material& m, f3 lv, f3 ll, float& pdf, float& reflectance)
{
    if (lv.z <= 0.0f || ll.z <= 0.0f)
    {
        pdf = 0.0f;
        reflectance = lv.z > 0.0f ? sheen_ltc_params(g, m.sheen_roughness, lv.z).b : 0.0f;
        return col(0.0f);
    }
    float phi = get_phi(lv);
    f3 l_std = rotate_vector(ll, mk(0.0f, 0.0f, 1.0f), -phi);
    Color AiBiRi = sheen_ltc_params(g, m.sheen_roughness, lv.z);
    float Do = eval_ltc(l_std, AiBiRi);
    pdf = Do;
    reflectance = AiBiRi.b;
    return m.sheen_color * AiBiRi.b * Do / ll.z;
}

PT_HD float sheen_pdf(const BsdfGlobals& g, const Material& m, f3 lv, f3 ll)
{
    if (lv.z <= 0.0f || ll.z <= 0.0f)
        return 0.0f;
    float phi = get_phi(lv);
    f3 l_std = rotate_vector(ll, mk(0.0f, 0.0f, 1.0f), -phi);
    Color AiBiRi = sheen_ltc_params(g, m.sheen_roughness, lv.z);
    return eval_ltc(l_std, AiBiRi);
}

// Sampling.h:210-220
PT_HD f3 cosine_sample_z_up(Xorshift32& rng)
{
    float r1 = rng();
    float r2 = rng();
    float phi = TWO_PI_F * r1;
    float cos_theta = sqrtf(r2);
    float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    float s, co;
    sincos_fixed(phi, s, co);
    return normalize(mk(co * sin_theta, s * sin_theta, cos_theta));
}

// SheenLTC.h:146-175: a cosine sample mapped through the LTC matrix M.
PT_HD f3 sheen_sample(const BsdfGlobals& g, const Material& m, f3 lv, Xorshift32& rng)
{
    f3 cs = cosine_sample_z_up(rng);
    Color AiBiRi = sheen_ltc_params(g, m.sheen_roughness, lv.z);
    float Ai_inv = 1.0f / AiBiRi.r;
    float Bi = AiBiRi.g;
    f3 d = normalize(mk(cs.x * Ai_inv - cs.z * Bi * Ai_inv, cs.y * Ai_inv, cs.z));
    return rotate_vector(d, mk(0.0f, 0.0f, 1.0f), get_phi(lv));
}

// ---------------------------------------------------------------------------
// Lobe weights and sampling probabilities

// ---------------------------------------------------------------------------
// Glass (specular transmission) and diffuse transmission lobes
// ---------------------------------------------------------------------------
// Sampling.h:160-171 (no total-internal-reflection check: NaN then).
PT_HD f3 refract_ray(f3 d, f3 n, float relative_eta)
{
    float NoI = dot(d, n);
    float sin_theta_i2 = 1.0f - NoI * NoI;
    float root = 1.0f - sin_theta_i2 / (relative_eta * relative_eta);
    float cos_theta_t = sqrtf(root);
    return vdiv(-1.0f * d, relative_eta) + (NoI / relative_eta - cos_theta_t) * n;
}

// MaterialUtils.h:32-57 (Kulla & Conty 2017 remapping of a thin interface).
PT_HD float thin_walled_roughness(bool thin_walled, float roughness, float relative_eta)
{
    if (!thin_walled)
        return roughness;
    float sq = (relative_eta - 0.5f) * (relative_eta - 0.5f);
    float cube = relative_eta * relative_eta * relative_eta;
    float remapped = roughness * sqrtf(3.7f * (relative_eta - 1.0f) * sq / cube);
    return clampf_(0.0f, 1.0f, remapped / 1.39f);
}

// MicrofacetRegularization.h:60-96: tau of the refraction lobe uses the
// consistent parameterisation of sample_number + 1.
PT_HD float regularize_refraction(const BsdfGlobals& g, uint32_t mode, float roughness, float path_roughness,
                                  float eta_i, float eta_t)
{
    if (!g.regularize || mode == NO_REGULARIZATION)
        return roughness;
    float path_tau = g.consistent_tau_refract / maxf_(path_roughness * path_roughness, 1.0e-8f);
    float d = eta_i - eta_t, mx = maxf_(eta_i, eta_t);
    float reg = sqrtf(sqrtf(1.0f / (path_tau * PI_F * (d * d) / (4.0f * (mx * mx)))));
    return maxf_(g.min_roughness, maxf_(roughness, reg));
}

// MicrofacetRegularization.h:98-146: 3/4 refraction + 1/4 reflection.
PT_HD float regularize_mix(const BsdfGlobals& g, uint32_t mode, float roughness, float path_roughness, float eta_i,
                           float eta_t)
{
    if (!g.regularize || mode == NO_REGULARIZATION)
        return roughness;
    float path_tau = g.consistent_tau / maxf_(path_roughness * path_roughness, 1.0e-8f);
    float reflection = sqrtf(sqrtf(1.0f / (path_tau * PI_F)));
    if (eta_i == eta_t)
        return reflection;
    float d = eta_i - eta_t, mx = maxf_(eta_i, eta_t);
    float refraction = sqrtf(sqrtf(1.0f / (path_tau * PI_F * (d * d) / (4.0f * (mx * mx)))));
    return maxf_(g.min_roughness, maxf_(roughness, refraction * 0.75f + reflection * 0.25f));
}

// One IOR knot of the correction exponent: the first roughness segment with
// roughness <= r_max gives a constant or lerp(a, b, (roughness - r0) / 0.1)
// with r0 the bound of the segment before.
PT_HD float glass_exponent_knot(const float* seg, float roughness)
{
    for (int k = 0; k < 11; k++)
    {
        if (roughness <= seg[4 * k])
        {
            if (seg[4 * k + 1] == 0.0f)
                return seg[4 * k + 2];
            float r0 = seg[4 * (k - 1)];
            return lerpf_(seg[4 * k + 2], seg[4 * k + 3], (roughness - r0) / 0.1f);
        }
    }
    return 2.5f;
}

// GGX_glass_energy_compensation_get_correction_exponent
// (MicrofacetEnergyCompensation.h:101-631): the exponent of the cos_theta
// axis of the glass LUTs, interpolated between the two IOR knots around
// relative_eta.
// The ten relative IORs the exponent is tabulated at (a switch, not an array:
// a dynamically indexed local array would live in scratch memory).
PT_HD float glass_exponent_ior_knot(int k)
{
    switch (k)
    {
    case 0: return 1.01f;
    case 1: return 1.02f;
    case 2: return 1.03f;
    case 3: return 1.1f;
    case 4: return 1.2f;
    case 5: return 1.4f;
    case 6: return 1.5f;
    case 7: return 2.0f;
    case 8: return 2.4f;
    default: return 3.0f;
    }
}

PT_HD float glass_correction_exponent(const float* tab, float roughness, float relative_eta)
{
    if ((roughness < 1.0e-10f && roughness > -1.0e-10f) || fabsf(1.0f - relative_eta) < 1.0e-3f)
        return 2.5f;
    float lower_bound = 1.01f, lower = 2.5f;
    for (int k = 0; k < 9; k++)
    {
        if (relative_eta > glass_exponent_ior_knot(k) && relative_eta <= glass_exponent_ior_knot(k + 1))
        {
            lower_bound = glass_exponent_ior_knot(k);
            lower = glass_exponent_knot(tab + 44 * k, roughness);
            break;
        }
    }
    int hk = 9;
    for (int k = 0; k < 9; k++)
    {
        if (relative_eta <= glass_exponent_ior_knot(k))
        {
            hk = k;
            break;
        }
    }
    float higher_bound = glass_exponent_ior_knot(hk);
    float higher = glass_exponent_knot(tab + 44 * (10 + hk), roughness);
    if (higher_bound == lower_bound)
        return lower;
    return lerpf_(lower, higher, (relative_eta - lower_bound) / (higher_bound - lower_bound));
}

// get_GGX_energy_compensation_dielectrics (MicrofacetEnergyCompensation.h:633-703),
// roughness threshold 0 and no bounce limit (BSDFsData.h:63,71); thin-film
// materials are refused, so the thin-film lerp is the identity.
PT_HD float glass_energy_compensation(const BsdfGlobals& g, const Material& m, float roughness, bool inside,
                                      float eta_t, float eta_i, float relative_eta, float NoV)
{
    bool smooth_enough = roughness <= 0.0f;
    if (!(m.flags & MAT_GLASS_EC) || smooth_enough)
        return 1.0f;
    float compensation = 1.0f;
    if (m.thin_film < 1.0f)
    {
        const bool thin = (m.flags & MAT_THIN_WALLED) != 0;
        float relative_eta_for_correction = inside ? 1.0f / relative_eta : relative_eta;
        float exponent = 2.5f;
        if (!thin)
            exponent = glass_correction_exponent(g.glass_exponent, roughness, relative_eta_for_correction);
        float view = pow_fixed(maxf_(1.0e-3f, NoV), 1.0f / exponent);
        float F0 = F0_from_eta(eta_t, eta_i);
        float F0_remapped = sqrtf(sqrtf(F0));
        if (thin)
            compensation = lut3d_fetch(g.thin_glass_ess, 32, 32, 96, view, roughness, F0_remapped);
        else
            compensation = lut3d_fetch(inside ? g.glass_inv_ess : g.glass_ess, 256, 16, 128, view, roughness,
                                       F0_remapped);
        compensation = lerpf_(compensation, 1.0f, m.thin_film * roughness);
    }
    return compensation;
}

// principled_beer_absorption (Principled.h:332-357): absorption of the medium
// the ray travelled in (the incident one), with the distance floored at 1e-6
// in the volume state itself.
PT_HD Color beer_absorption(const BsdfGlobals& g, VolumeState& v)
{
    const Material& medium = g.mats[v.incident];
    const Color a = medium.absorption_color;
    if (!is_white(a))
    {
        v.distance = maxf_(v.distance, 1.0e-6f);
        Color k = col(log_fixed(a.r), log_fixed(a.g), log_fixed(a.b)) / medium.absorption_at_distance;
        Color e = k * v.distance;
        return col(exp_fixed(e.r), exp_fixed(e.g), exp_fixed(e.b));
    }
    return col(1.0f);
}

// Walter et al. 2007 refraction lobe (Microfacet.h:288-352).
PT_HD Color ggx_eval_refract(const Material& m, float roughness, float relative_eta, Color F, f3 lv, f3 ll, f3 lh,
                             float& pdf, uint32_t info)
{
    float NoL = ll.z, NoV = lv.z;
    float HoL = dot(ll, lh), HoV = dot(lv, lh);
    Color color;
    if (is_perfectly_smooth(roughness))
    {
        bool perfect = dot(refract_ray(lv, mk(0.0f, 0.0f, 1.0f), relative_eta), ll) > DELTA_ALIGN_THRESHOLD;
        if (info == FROM_GLASS_REFRACT && perfect)
        {
            color = col(DELTA_HIGH_VALUE) * (col(1.0f) - F) * m.base_color;
            color = color / fabsf(NoL);
            pdf = DELTA_HIGH_VALUE;
        }
        else
        {
            color = col(0.0f);
            pdf = 0.0f;
        }
    }
    else
    {
        float dot_prod = HoL + HoV / relative_eta;
        float dot_prod2 = dot_prod * dot_prod;
        float denom = dot_prod2 * NoL * NoV;
        float ax, ay;
        get_alphas(roughness, m.anisotropy, ax, ay);
        float D = ggx_d(ax, ay, lh);
        float G1_V = 1.0f / (1.0f + smith_lambda(ax, ay, lv));
        float G1_L = 1.0f / (1.0f + smith_lambda(ax, ay, ll));
        float G2 = G1_V * G1_L;
        float dwm_dwi = fabsf(HoL) / dot_prod2;
        float D_pdf = G1_V / fabsf(NoV) * D * fabsf(HoV);
        pdf = dwm_dwi * D_pdf;
        color = m.base_color * D * (col(1.0f) - F) * G2 * fabsf(HoL * HoV / denom);
    }
    color = color / (relative_eta * relative_eta);
    return color;
}

PT_HD float ggx_pdf_refract(const Material& m, float roughness, float relative_eta, f3 lv, f3 ll, f3 lh, uint32_t info)
{
    float NoV = lv.z;
    float HoL = dot(ll, lh), HoV = dot(lv, lh);
    if (is_perfectly_smooth(roughness))
    {
        bool perfect = dot(refract_ray(lv, mk(0.0f, 0.0f, 1.0f), relative_eta), ll) > DELTA_ALIGN_THRESHOLD;
        return info == FROM_GLASS_REFRACT && perfect ? DELTA_HIGH_VALUE : 0.0f;
    }
    float dot_prod = HoL + HoV / relative_eta;
    float dot_prod2 = dot_prod * dot_prod;
    float ax, ay;
    get_alphas(roughness, m.anisotropy, ax, ay);
    float D = ggx_d(ax, ay, lh);
    float G1_V = 1.0f / (1.0f + smith_lambda(ax, ay, lv));
    float dwm_dwi = fabsf(HoL) / dot_prod2;
    float D_pdf = G1_V / fabsf(NoV) * D * fabsf(HoV);
    return dwm_dwi * D_pdf;
}

// IORs on both sides of the interface (Principled.h:374-380): the materials
// of the incident / outgoing media (1 for the air) at the sampled wavelength.
PT_HD float medium_ior(const BsdfGlobals& g, int32_t medium)
{
    return (uint32_t)medium == VOL_AIR ? 1.0f : g.mats[medium].ior;
}

PT_HD void glass_etas(const BsdfGlobals& g, const Material& m, const VolumeState& v, float& eta_i, float& eta_t,
                      float& relative_eta)
{
    eta_i = medium_ior(g, v.incident);
    eta_t = medium_ior(g, v.outgoing);
    eta_i = dispersion_ior(m.dispersion_abbe_number, m.dispersion_scale, eta_i, fabsf(v.wavelength));
    eta_t = dispersion_ior(m.dispersion_abbe_number, m.dispersion_scale, eta_t, fabsf(v.wavelength));
    relative_eta = eta_t / eta_i;
    // a relative IOR of exactly 1 with opposite view and light gives a null half vector
    if (fabsf(relative_eta - 1.0f) < 1.0e-5f)
        relative_eta = 1.0f + 1.0e-5f;
}

// Fresnel reflectance of the glass interface (lerped with the thin-film one)
// and the probability of sampling the reflection (its luminance). A thin
// interface without film adds the inter-reflections between its two sides
// (PBR book 4th ed., dielectric BSDF).
PT_HD Color glass_fresnel(const Material& m, bool thin_walled, float scaled_roughness, float eta_i, float relative_eta,
                          float HoV, float& reflect_proba)
{
    Color F_thin_film = thin_film_fresnel(m, eta_i, HoV);
    Color F_plain = col(0.0f);
    if (m.thin_film < 1.0f)
        F_plain = col(fresnel_dielectric(HoV, relative_eta));
    Color F = lerpc(F_plain, F_thin_film, m.thin_film);
    if (thin_walled && F.r < 1.0f && m.thin_film == 0.0f && scaled_roughness < 0.1f)
        F = F + col((1.0f - F.r) * (1.0f - F.r) * F.r / (1.0f - F.r * F.r));
    reflect_proba = luminance(F);
    return F;
}

// Generalised half vector of the glass lobe (Principled.h:410-443), flipped
// into the upper hemisphere.
PT_HD f3 glass_half_vector(const BsdfGlobals& g, bool reflecting, bool thin_walled, float scaled_roughness,
                           float relative_eta, f3 lv, f3 ll)
{
    f3 lh;
    if (scaled_roughness <= ROUGHNESS_CLAMP && !g.regularize)
        lh = mk(0.0f, 0.0f, 1.0f);
    else
    {
        if (reflecting)
            lh = ll + lv;
        else if (thin_walled)
            lh = mk(ll.x, ll.y, -ll.z) + lv;
        else
            lh = mk(ll.x * relative_eta, ll.y * relative_eta, ll.z * relative_eta) + lv;
        lh = normalize(lh);
    }
    if (lh.z < 0.0f)
        lh = -1.0f * lh;
    return lh;
}

// principled_glass_eval (Principled.h:359-558), in the rotated frame.
PT_HD Color glass_eval(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, f3 lv, f3 ll, float& pdf)
{
    pdf = 0.0f;
    float NoV = lv.z, NoL = ll.z;
    if (fabsf(NoL) < 1.0e-8f)
        return col(0.0f);
    bool reflecting = NoL * NoV > 0.0f;
    VolumeState& v = *c.vs;
    float eta_i, eta_t, relative_eta;
    glass_etas(g, m, v, eta_i, eta_t, relative_eta);
    const bool thin = (m.flags & MAT_THIN_WALLED) != 0;
    float scaled = thin_walled_roughness(thin, m.roughness, relative_eta);
    f3 lh = glass_half_vector(g, reflecting, thin, scaled, relative_eta, lv, ll);
    float HoL = dot(ll, lh), HoV = dot(lv, lh);
    if (HoL * NoL < 0.0f || HoV * NoV < 0.0f)
        return col(0.0f);
    float reflect_proba;
    Color F = glass_fresnel(m, thin, scaled, eta_i, relative_eta, HoV, reflect_proba);
    Color color;
    if (reflecting)
    {
        float rr = scaled;
        if (g.regularize && c.reg == REGULARIZATION_MIS)
            rr = regularize_mix(g, c.reg, scaled, c.path_roughness, eta_i, eta_t);
        else if (g.regularize && c.reg == REGULARIZATION_CLASSIC)
            rr = regularize_reflection(g, c.reg, scaled, c.path_roughness);
        DeltaSampled ds = delta_reflection_sampled(m, scaled, m.anisotropy, c.info);
        color = ggx_eval_reflect_ss(rr, m.anisotropy, F, lv, ll, lh, pdf, ds);
        float compensation = glass_energy_compensation(g, m, m.roughness, v.inside != 0, eta_t, eta_i, relative_eta,
                                                       lv.z);
        color = color / compensation;
        pdf *= reflect_proba;
    }
    else
    {
        float rr = scaled;
        if (g.regularize && c.reg == REGULARIZATION_MIS)
            rr = regularize_mix(g, c.reg, scaled, c.path_roughness, eta_i, eta_t);
        else if (g.regularize && c.reg == REGULARIZATION_CLASSIC)
            rr = regularize_refraction(g, c.reg, scaled, c.path_roughness, eta_i, eta_t);
        color = ggx_eval_refract(m, rr, relative_eta, F, lv, ll, lh, pdf, c.info);
        pdf *= 1.0f - reflect_proba;
        float compensation = glass_energy_compensation(g, m, rr, v.inside != 0, eta_t, eta_i, relative_eta, lv.z);
        color = color / compensation;
        if (thin)
            color = color * m.base_color; // both sides of the thin interface at once
        if (thin && c.update)
            vol_pop(v, v.inside != 0);
        else if ((uint32_t)v.incident != VOL_AIR)
        {
            color = color * beer_absorption(g, v);
            if (c.update)
            {
                v.distance = 0.0f;
                v.reset = 1;
                if (v.inside)
                    vol_pop(v, true);
            }
        }
    }
    return color;
}

// principled_glass_pdf (Principled.h:560-724)
PT_HD float glass_pdf(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, f3 lv, f3 ll)
{
    float NoV = lv.z, NoL = ll.z;
    if (fabsf(NoL) < 1.0e-8f)
        return 0.0f;
    bool reflecting = NoL * NoV > 0.0f;
    const VolumeState& v = *c.vs;
    float eta_i, eta_t, relative_eta;
    glass_etas(g, m, v, eta_i, eta_t, relative_eta);
    const bool thin = (m.flags & MAT_THIN_WALLED) != 0;
    float scaled = thin_walled_roughness(thin, m.roughness, relative_eta);
    f3 lh = glass_half_vector(g, reflecting, thin, scaled, relative_eta, lv, ll);
    float HoL = dot(ll, lh), HoV = dot(lv, lh);
    if (HoL * NoL < 0.0f || HoV * NoV < 0.0f)
        return 0.0f;
    float reflect_proba;
    glass_fresnel(m, thin, scaled, eta_i, relative_eta, HoV, reflect_proba);
    float pdf;
    if (reflecting)
    {
        float rr = scaled;
        if (g.regularize && c.reg == REGULARIZATION_MIS)
            rr = regularize_mix(g, c.reg, scaled, c.path_roughness, eta_i, eta_t);
        else if (g.regularize && c.reg == REGULARIZATION_CLASSIC)
            rr = regularize_reflection(g, c.reg, scaled, c.path_roughness);
        DeltaSampled ds = delta_reflection_sampled(m, scaled, m.anisotropy, c.info);
        pdf = ggx_pdf_reflect(rr, m.anisotropy, lv, ll, lh, ds);
        pdf *= reflect_proba;
    }
    else
    {
        float rr = scaled;
        if (g.regularize && c.reg == REGULARIZATION_MIS)
            rr = regularize_mix(g, c.reg, scaled, c.path_roughness, eta_i, eta_t);
        else if (g.regularize && c.reg == REGULARIZATION_CLASSIC)
            rr = regularize_refraction(g, c.reg, scaled, c.path_roughness, eta_i, eta_t);
        pdf = ggx_pdf_refract(m, rr, relative_eta, lv, ll, lh, c.info);
        pdf *= 1.0f - reflect_proba;
    }
    return pdf;
}

// principled_glass_sample (Principled.h:726-830): a microfacet normal from the
// visible normals, then reflection with the Fresnel probability, else
// refraction (straight through a thin interface).
PT_HD f3 glass_sample(const BsdfGlobals& g, const Material& m, BsdfCtx& c, f3 lv, Xorshift32& rng)
{
    VolumeState& v = *c.vs;
    float eta_i, eta_t, relative_eta;
    glass_etas(g, m, v, eta_i, eta_t, relative_eta);
    const bool thin = (m.flags & MAT_THIN_WALLED) != 0;
    float scaled = thin_walled_roughness(thin, m.roughness, relative_eta);
    if (g.regularize && c.reg == REGULARIZATION_MIS)
        scaled = regularize_mix(g, c.reg, scaled, c.path_roughness, eta_i, eta_t);
    float ax, ay;
    get_alphas(scaled, m.anisotropy, ax, ay);
    f3 mn = (ax <= ROUGHNESS_CLAMP && ay <= ROUGHNESS_CLAMP) ? mk(0.0f, 0.0f, 1.0f) : ggx_vndf_sample(lv, ax, ay, rng);
    float HoV = dot(lv, mn);
    float reflect_proba;
    glass_fresnel(m, thin, scaled, eta_i, relative_eta, HoV, reflect_proba);
    float rand_1 = rng();
    if (rand_1 < reflect_proba)
    {
        c.info = FROM_GLASS_REFLECT;
        if (c.update)
            vol_pop(v, false);
        return reflect_ray(lv, mn);
    }
    c.info = FROM_GLASS_REFRACT;
    if (dot(mn, lv) < 0.0f)
        mn = -1.0f * mn;
    if (thin)
    {
        f3 reflected = reflect_ray(lv, mn);
        reflected.z *= -1.0f;
        if (c.update)
            vol_pop(v, false);
        return reflected;
    }
    return refract_ray(lv, mn, relative_eta);
}

// principled_diffuse_transmission_eval / _pdf (Principled.h:832-873).
PT_HD Color dtrans_eval(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, f3 lv, f3 ll, float& pdf)
{
    pdf = 0.0f;
    if (lv.z * ll.z > 0.0f)
        return col(0.0f);
    Color color = m.base_color * INV_PI_F;
    VolumeState& v = *c.vs;
    if ((uint32_t)v.incident != VOL_AIR)
    {
        color = color * beer_absorption(g, v);
        if (c.update)
        {
            v.distance = 0.0f;
            v.reset = 1;
            if (v.inside)
                vol_pop(v, true);
        }
    }
    pdf = fabsf(ll.z * INV_PI_F);
    return color;
}

PT_HD float dtrans_pdf(f3 lv, f3 ll)
{
    if (lv.z * ll.z > 0.0f)
        return 0.0f;
    return fabsf(ll.z * INV_PI_F);
}

// ---------------------------------------------------------------------------
struct LobeWeights {
    float coat, sheen, metal1, metal2, specular, diffuse, glass, dtrans;
};

// Principled.h:1515-1554. Inside an object only the two transmission lobes
// remain (the path is leaving the medium through its boundary).
PT_HD LobeWeights lobe_weights(const Material& m, bool outside_object)
{
    const float outside = outside_object ? 1.0f : 0.0f;
    LobeWeights w;
    w.coat = m.coat * outside;
    w.sheen = m.sheen * outside;
    float metallic = m.metallic;
    w.metal1 = metallic * outside;
    w.metal2 = metallic * outside;
    w.metal1 = lerpf_(w.metal1, 0.0f, m.second_roughness_weight);
    w.metal2 = lerpf_(0.0f, w.metal2, m.second_roughness_weight);
    float st = m.specular_transmission, dt = m.diffuse_transmission;
    w.glass = !outside_object ? (1.0f - dt) : (1.0f - metallic) * (1.0f - dt) * st;
    w.dtrans = !outside_object ? dt : (1.0f - metallic) * dt;
    w.specular = (1.0f - metallic) * (1.0f - st * (1.0f - dt)) * m.specular * outside;
    w.diffuse = (1.0f - metallic) * (1.0f - st) * (1.0f - dt) * outside;
    return w;
}

// Principled.h:1556-1618 (no Fresnel-based lobe sampling: both options off).
PT_HD LobeWeights lobe_probas(const LobeWeights& w)
{
    float nf = 1.0f / (w.coat + w.sheen + w.metal1 + w.metal2 + w.specular + w.diffuse + w.glass + w.dtrans);
    LobeWeights p;
    p.coat = w.coat * nf;
    p.sheen = w.sheen * nf;
    p.metal1 = w.metal1 * nf;
    p.metal2 = w.metal2 * nf;
    p.specular = w.specular * nf;
    p.diffuse = w.diffuse * nf;
    p.glass = w.glass * nf;
    p.dtrans = w.dtrans * nf;
    return p;
}

// The three local frames of one evaluation (Principled.h:1633-1645).
struct LocalFrames {
    f3 lv, ll, lh;    // unrotated ONB
    f3 lvr, llr, lhr; // ONB rotated by anisotropy_rotation * pi
};

PT_HD LocalFrames local_frames(const Material& m, const BsdfCtx& c)
{
    LocalFrames f;
    f3 T, B;
    build_onb(c.sn, T, B);
    f.lv = world_to_local(T, B, c.sn, c.view);
    f.ll = world_to_local(T, B, c.sn, c.to_light);
    f.lh = normalize(f.lv + f.ll);
    f3 TR, BR;
    build_rotated_onb(c.sn, TR, BR, m.anisotropy_rotation * PI_F);
    f.lvr = world_to_local(TR, BR, c.sn, c.view);
    f.llr = world_to_local(TR, BR, c.sn, c.to_light);
    f.lhr = normalize(f.lvr + f.llr);
    return f;
}

// ---------------------------------------------------------------------------
// Layers (Principled.h:955-1512), evaluated top to bottom; `layers` is the
// energy left for the layers below.
// ---------------------------------------------------------------------------
// Principled.h:955-1058
PT_HD Color eval_coat_layer(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, const LocalFrames& f,
                            float incident_ior, bool refracting, float weight, float proba, Color& layers,
                            float& pdf)
{
    float coat_ior = m.coat_ior;
    if (weight > 0.0f && ((f.lv.z > 0.0f && f.ll.z > 0.0f) || refracting))
    {
        float cpdf = 0.0f;
        Color contribution = col(0.0f);
        if (!refracting)
        {
            contribution = coat_eval(g, m, c, f.lv, f.ll, f.lh, incident_ior, cpdf);
            contribution = contribution * weight;
            contribution = contribution * layers;
        }
        pdf += cpdf * proba;
        Color below = col(1.0f);
        below = below * (1.0f - fresnel_dielectric(fabsf(f.ll.z), incident_ior, coat_ior));
        float view_dir_fresnel = fresnel_dielectric(fabsf(f.lv.z), incident_ior, coat_ior);
        below = below * (1.0f - view_dir_fresnel);
        if (!is_white(m.coat_medium_absorption))
        {
            float in_angle = maxf_(1.0e-6f, sqrtf(1.0f - (1.0f - f.ll.z * f.ll.z) / (coat_ior * coat_ior)));
            float out_angle = maxf_(1.0e-6f, sqrtf(1.0f - (1.0f - f.lv.z * f.lv.z) / (coat_ior * coat_ior)));
            float distance = 1.0f / in_angle + 1.0f / out_angle;
            const Color a = m.coat_medium_absorption;
            Color p = col(pow_fixed(sqrtf(a.r), distance), pow_fixed(sqrtf(a.g), distance),
                          pow_fixed(sqrtf(a.b), distance));
            Color e = (col(0.0f) - (col(1.0f) - p)) * m.coat_medium_thickness;
            below = below * col(exp_fixed(e.r), exp_fixed(e.g), exp_fixed(e.b));
        }
        below = below * coat_darkening(m, coat_ior / incident_ior, view_dir_fresnel);
        below = lerpc(col(1.0f), below, m.coat);
        layers = layers * below;
        return contribution;
    }
    return col(0.0f);
}

PT_HD float pdf_coat_layer(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, const LocalFrames& f,
                           bool refracting, float weight, float proba)
{
    if (weight > 0.0f && ((f.lv.z > 0.0f && f.ll.z > 0.0f) || refracting))
    {
        float cpdf = 0.0f;
        if (!refracting)
            cpdf = coat_pdf(g, m, c, f.lv, f.ll, f.lh);
        return cpdf * proba;
    }
    return 0.0f;
}

// Principled.h:1088-1113
PT_HD Color eval_sheen_layer(const BsdfGlobals& g, const Material& m, const LocalFrames& f, bool refracting,
                             float weight, float proba, Color& layers, float& pdf)
{
    if ((weight > 0.0f && f.lv.z > 0.0f && f.ll.z > 0.0f) || refracting)
    {
        float reflectance, spdf;
        Color contribution = sheen_eval(g, m, f.lv, f.ll, spdf, reflectance);
        contribution = contribution * weight;
        contribution = contribution * layers;
        pdf += spdf * proba;
        layers = layers * (1.0f - m.sheen * reflectance);
        return contribution;
    }
    return col(0.0f);
}

PT_HD float pdf_sheen_layer(const BsdfGlobals& g, const Material& m, const LocalFrames& f, bool refracting,
                            float weight, float proba)
{
    if ((weight > 0.0f && f.lv.z > 0.0f && f.ll.z > 0.0f) || refracting)
        return sheen_pdf(g, m, f.lv, f.ll) * proba;
    return 0.0f;
}

// Principled.h:1129-1157 (the metal lobe has nothing below it).
PT_HD Color eval_metal_layer(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, const LocalFrames& f,
                             float roughness, float incident_ior, float weight, float proba, const Color& layers,
                             float& pdf)
{
    if (weight > 0.0f && f.lvr.z > 0.0f && f.llr.z > 0.0f)
    {
        float mpdf = 0.0f;
        Color contribution = metallic_eval(g, m, c, roughness, m.anisotropy, incident_ior, f.lvr, f.llr, f.lhr, mpdf);
        contribution = contribution * weight;
        contribution = contribution * layers;
        pdf += mpdf * proba;
        return contribution;
    }
    return col(0.0f);
}

PT_HD float pdf_metal_layer(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, const LocalFrames& f,
                            float roughness, float weight, float proba)
{
    if (weight > 0.0f && f.lvr.z > 0.0f && f.llr.z > 0.0f)
        return metallic_pdf(g, m, c, roughness, m.anisotropy, f.lvr, f.llr, f.lhr) * proba;
    return 0.0f;
}

// Principled.h:1305-1405, in the rotated frame.
PT_HD Color eval_specular_layer(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, const LocalFrames& f,
                                float incident_ior, float weight, bool refracting, float proba, Color& layers,
                                float& pdf)
{
    if (weight > 0.0f && ((f.lvr.z > 0.0f && f.llr.z > 0.0f) || (refracting && c.info == FROM_DIFFUSE_TRANSMISSION)))
    {
        float relative_ior = specular_relative_ior(m, incident_ior);
        float spdf = 0.0f;
        Color contribution = specular_eval(g, m, c, relative_ior, f.lvr, f.llr, f.lhr, spdf);
        contribution = contribution * lerpc(col(1.0f), m.specular_tint * m.specular_color, m.specular);
        contribution = contribution * weight;
        contribution = contribution * layers;
        Color below = col(1.0f);
        Color light_dir_fresnel = specular_fresnel(m, relative_ior, fabsf(f.llr.z));
        light_dir_fresnel = light_dir_fresnel * (1.0f - m.diffuse_transmission);
        below = below * (col(1.0f) - light_dir_fresnel);
        Color view_dir_fresnel = specular_fresnel(m, relative_ior, fabsf(f.lvr.z));
        below = below * (col(1.0f) - view_dir_fresnel);
        below = below * specular_darkening(m, relative_ior, luminance(view_dir_fresnel));
        below = lerpc(col(1.0f), below, m.specular);
        layers = layers * below;
        pdf += spdf * proba;
        return contribution;
    }
    return col(0.0f);
}

PT_HD float pdf_specular_layer(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, const LocalFrames& f,
                               float weight, bool refracting, float proba)
{
    if (weight > 0.0f && ((f.lvr.z > 0.0f && f.llr.z > 0.0f) || (refracting && c.info == FROM_DIFFUSE_TRANSMISSION)))
        return specular_pdf(g, m, c, f.lvr, f.llr, f.lhr) * proba;
    return 0.0f;
}

// Principled.h:1436-1469 (Lambertian diffuse, Lambertian.h:14-38).
PT_HD Color eval_diffuse_layer(const Material& m, const LocalFrames& f, float weight, float proba, const Color& layers,
                               float& pdf)
{
    if (weight > 0.0f && f.lv.z > 0.0f && f.ll.z > 0.0f)
    {
        float dpdf = 0.0f;
        Color contribution = col(0.0f);
        if (f.ll.z > 0.0f)
        {
            dpdf = f.ll.z * INV_PI_F;
            contribution = m.base_color * INV_PI_F;
        }
        contribution = contribution * weight;
        contribution = contribution * layers;
        pdf += dpdf * proba;
        return contribution;
    }
    return col(0.0f);
}

PT_HD float pdf_diffuse_layer(const LocalFrames& f, float weight, float proba)
{
    if (weight > 0.0f && f.lv.z > 0.0f && f.ll.z > 0.0f)
        return (f.ll.z > 0.0f ? f.ll.z * INV_PI_F : 0.0f) * proba;
    return 0.0f;
}

// ---------------------------------------------------------------------------
// Principled BSDF (Principled.h:1620-1939)
// ---------------------------------------------------------------------------
// Principled.h:1556-1618 and :1620-1715
PT_HD Color principled_eval(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, float& pdf)
{
    pdf = 0.0f;
    const VolumeState& v = *c.vs;
    bool outside = !v.inside;
    bool refracting = dot(c.sn, c.to_light) < 0.0f && outside;
    LocalFrames f = local_frames(m, c);
    const float incident_ior = medium_ior(g, v.incident);
    LobeWeights w = lobe_weights(m, outside);
    LobeWeights p = lobe_probas(w);
    Color layers = col(1.0f);
    Color final_color = col(0.0f);
    final_color = final_color + eval_coat_layer(g, m, c, f, incident_ior, refracting, w.coat, p.coat, layers, pdf);
    final_color = final_color + eval_sheen_layer(g, m, f, refracting, w.sheen, p.sheen, layers, pdf);
    final_color = final_color + eval_metal_layer(g, m, c, f, m.roughness, incident_ior, w.metal1 * (float)!refracting,
                                                 p.metal1, layers, pdf);
    final_color = final_color + eval_metal_layer(g, m, c, f, m.second_roughness, incident_ior,
                                                 w.metal2 * (float)!refracting, p.metal2, layers, pdf);
    // glass layer (Principled.h:1177-1200), rotated frame
    if (w.glass > 0.0f)
    {
        float gpdf = 0.0f;
        Color contribution = glass_eval(g, m, c, f.lvr, f.llr, gpdf);
        contribution = contribution * w.glass;
        contribution = contribution * layers;
        pdf += gpdf * p.glass;
        final_color = final_color + contribution;
    }
    // glossy base: specular layer over the diffuse layer (Principled.h:1474-1492)
    Color glossy = col(0.0f);
    glossy = glossy + eval_specular_layer(g, m, c, f, incident_ior, w.specular, refracting, p.specular, layers, pdf);
    glossy = glossy + eval_diffuse_layer(m, f, w.diffuse * (float)!refracting, p.diffuse, layers, pdf);
    final_color = final_color + glossy / energy_compensation_glossy_base(g, m, incident_ior, f.lv.z);
    // diffuse transmission layer (Principled.h:1219-1241)
    if (w.dtrans > 0.0f && f.ll.z < 0.0f)
    {
        float dpdf = 0.0f;
        Color contribution = dtrans_eval(g, m, c, f.lv, f.ll, dpdf);
        contribution = contribution * w.dtrans;
        contribution = contribution * layers;
        pdf += dpdf * p.dtrans;
        final_color = final_color + contribution;
    }
    return final_color / energy_compensation_clearcoat(g, m, incident_ior, f.lv.z);
}

// Principled.h:1717-1800
PT_HD float principled_pdf(const BsdfGlobals& g, const Material& m, const BsdfCtx& c)
{
    float pdf = 0.0f;
    const VolumeState& v = *c.vs;
    bool outside = !v.inside;
    bool refracting = dot(c.sn, c.to_light) < 0.0f && outside;
    LocalFrames f = local_frames(m, c);
    LobeWeights w = lobe_weights(m, outside);
    LobeWeights p = lobe_probas(w);
    pdf += pdf_coat_layer(g, m, c, f, refracting, w.coat, p.coat);
    pdf += pdf_sheen_layer(g, m, f, refracting, w.sheen, p.sheen);
    pdf += pdf_metal_layer(g, m, c, f, m.roughness, w.metal1 * (float)!refracting, p.metal1);
    pdf += pdf_metal_layer(g, m, c, f, m.second_roughness, w.metal2 * (float)!refracting, p.metal2);
    if (w.glass > 0.0f)
        pdf += glass_pdf(g, m, c, f.lvr, f.llr) * p.glass;
    float glossy = 0.0f;
    glossy += pdf_specular_layer(g, m, c, f, w.specular, refracting, p.specular);
    glossy += pdf_diffuse_layer(f, w.diffuse * (float)!refracting, p.diffuse);
    pdf += glossy;
    if (w.dtrans > 0.0f && f.ll.z < 0.0f)
        pdf += dtrans_pdf(f.lv, f.ll) * p.dtrans;
    return pdf;
}

// Principled.h:1802-1939. Returns the BSDF value of the sampled direction;
// pdf = 0 when the sample is invalid. `c.info` receives the sampled lobe
// (the sheen lobe leaves it as it was). With c.update, a reflection pops the
// material of the vertex from the interior stack (the path stays outside).
PT_HD Color principled_sample(const BsdfGlobals& g, const Material& m, BsdfCtx& c, f3& out_dir, float& pdf,
                              Xorshift32& rng)
{
    pdf = 0.0f;
    VolumeState& v = *c.vs;
    bool outside = !v.inside;
    LobeWeights w = lobe_weights(m, outside);
    LobeWeights p = lobe_probas(w);
    float cdf0 = p.coat;
    float cdf1 = cdf0 + p.sheen;
    float cdf2 = cdf1 + p.metal1;
    float cdf3 = cdf2 + p.metal2;
    float cdf4 = cdf3 + p.specular;
    float cdf5 = cdf4 + p.diffuse;
    float cdf6 = cdf5 + p.dtrans;
    float rand_1 = rng();
    bool sampling_dt = rand_1 > cdf5 && rand_1 < cdf6;
    bool sampling_glass = rand_1 > cdf6;
    if (c.update && !sampling_glass && !sampling_dt)
        vol_pop(v, false);
    f3 TR, BR;
    build_rotated_onb(c.sn, TR, BR, m.anisotropy_rotation * PI_F);
    f3 lvr = world_to_local(TR, BR, c.sn, c.view);
    if (rand_1 < cdf0)
    {
        f3 TC, BC;
        build_rotated_onb(c.sn, TC, BC, m.coat_anisotropy_rotation * PI_F);
        f3 lvc = world_to_local(TC, BC, c.sn, c.view);
        c.info = FROM_COAT;
        float rr = regularize_reflection(g, c.reg, m.coat_roughness, c.path_roughness);
        out_dir = local_to_world(TC, BC, c.sn, ggx_sample_reflection(rr, m.coat_anisotropy, lvc, rng));
    }
    else if (rand_1 < cdf1)
    {
        f3 T, B;
        build_onb(c.sn, T, B);
        f3 lv = world_to_local(T, B, c.sn, c.view);
        out_dir = local_to_world(T, B, c.sn, sheen_sample(g, m, lv, rng));
    }
    else if (rand_1 < cdf2)
    {
        c.info = FROM_FIRST_METAL;
        float rr = regularize_reflection(g, c.reg, m.roughness, c.path_roughness);
        out_dir = local_to_world(TR, BR, c.sn, ggx_sample_reflection(rr, m.anisotropy, lvr, rng));
    }
    else if (rand_1 < cdf3)
    {
        c.info = FROM_SECOND_METAL;
        float rr = regularize_reflection(g, c.reg, m.second_roughness, c.path_roughness);
        out_dir = local_to_world(TR, BR, c.sn, ggx_sample_reflection(rr, m.anisotropy, lvr, rng));
    }
    else if (rand_1 < cdf4)
    {
        c.info = FROM_SPECULAR;
        float rr = regularize_reflection(g, c.reg, m.roughness, c.path_roughness);
        out_dir = local_to_world(TR, BR, c.sn, ggx_sample_reflection(rr, m.anisotropy, lvr, rng));
    }
    else if (rand_1 < cdf5)
    {
        c.info = FROM_DIFFUSE;
        out_dir = cosine_sample_around_normal(c.sn, rng);
    }
    else if (rand_1 < cdf6)
    {
        c.info = FROM_DIFFUSE_TRANSMISSION;
        out_dir = cosine_sample_around_normal(-1.0f * c.sn, rng);
    }
    else
        out_dir = local_to_world(TR, BR, c.sn, glass_sample(g, m, c, lvr, rng));
    if (dot(out_dir, c.gn) < 0.0f && !sampling_glass && !sampling_dt)
        return col(0.0f);
    c.to_light = out_dir;
    return principled_eval(g, m, c, pdf);
}

// ---------------------------------------------------------------------------
// G-buffer normals: octahedral 2x12-bit snorm in 3 bytes + 1 padding byte
// (src/HostDeviceCommon/Packing.h:237-360, Cigolle et al. 2014).
// ---------------------------------------------------------------------------
PT_HD float sign_not_zero(float k) { return k >= 0.0f ? 1.0f : -1.0f; }

PT_HD uint32_t oct_pack(f3 v)
{
    float l1 = 1.0f / (fabsf(v.x) + fabsf(v.y) + fabsf(v.z));
    float ex = v.x * l1, ey = v.y * l1;
    if (v.z < 0.0f)
    {
        float nx = (1.0f - fabsf(ey)) * sign_not_zero(ex);
        float ny = (1.0f - fabsf(ex)) * sign_not_zero(ey);
        ex = nx;
        ey = ny;
    }
    float sx = roundf(clampf_(0.0f, 2.0f, ex + 1.0f) * 2047.0f);
    float sy = roundf(clampf_(0.0f, 2.0f, ey + 1.0f) * 2047.0f);
    float ux = sx / 16.0f;
    float t = floorf(sy / 256.0f);
    float uy = ((ux - floorf(ux)) * 256.0f) + t;
    float uz = sy - (t * 256.0f);
    uint32_t bx = (uint32_t)(uint8_t)ux, by = (uint32_t)(uint8_t)uy, bz = (uint32_t)(uint8_t)uz;
    return bx | (by << 8) | (bz << 16);
}

PT_HD f3 oct_unpack(uint32_t p)
{
    uint32_t bx = p & 0xffu, by = (p >> 8) & 0xffu, bz = (p >> 16) & 0xffu;
    float temp = (float)by / 16.0f;
    float sx = (float)bx * 16.0f + floorf(temp);
    float sy = (temp - floorf(temp)) * 256.0f * 16.0f + (float)bz;
    float x = clampf_(-1.0f, 1.0f, (sx / 2047.0f) - 1.0f);
    float y = clampf_(-1.0f, 1.0f, (sy / 2047.0f) - 1.0f);
    f3 v = mk(x, y, 1.0f - fabsf(x) - fabsf(y));
    if (v.z < 0.0f)
    {
        float tx = v.x, ty = v.y;
        v.x = (1.0f - fabsf(ty)) * sign_not_zero(tx);
        v.y = (1.0f - fabsf(tx)) * sign_not_zero(ty);
    }
    return normalize(v);
}

} // namespace pt
