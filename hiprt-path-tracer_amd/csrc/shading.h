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

// Unpacked material (after the reference's 8-bit packing round trip,
// MaterialPacked.h). 128 bytes.
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
    uint32_t flags; // bit 0: metallic energy compensation, bit 1: specular, bit 2: coat, bit 3: glass
    float pad[5];
};
static_assert(sizeof(Material) == 128, "Material must be 128 bytes");

constexpr uint32_t MAT_METALLIC_EC = 1u;
constexpr uint32_t MAT_SPECULAR_EC = 2u;

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
constexpr uint32_t RESTIR_FLAG_LOBE_MASK = 0x3fu << 1;

enum RegMode : uint32_t { NO_REGULARIZATION = 0, REGULARIZATION_CLASSIC = 1, REGULARIZATION_MIS = 2 };

// Per-pass constants of the BSDF (BSDFsData.h + MicrofacetRegularizationSettings.h).
struct BsdfGlobals {
    const float* ggx_conductor_ess; // 128 x 128, row 0 = roughness 1
    int32_t lut_size;
    float consistent_tau;           // MicrofacetRegularization::consistent_tau(tau_0, sample_number)
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

// MaterialUtils.h:143-175 (metal part; coat/glass lobes are not built yet).
PT_HD DeltaSampled delta_reflection_sampled(const Material& m, float roughness, float anisotropy, uint32_t info)
{
    if (!is_perfectly_smooth(roughness))
        return NOT_SPECULAR;
    bool matching = fabsf(anisotropy - m.anisotropy) < 1.0e-3f;
    bool first = info == FROM_FIRST_METAL && is_perfectly_smooth(m.roughness) && matching;
    bool second = info == FROM_SECOND_METAL && is_perfectly_smooth(m.second_roughness) && matching;
    if (first || second)
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
};

// Principled.h:125-153
PT_HD Color metallic_eval(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, float roughness, float anisotropy,
                          f3 lv, f3 ll, f3 lh, float& pdf)
{
    float rr = regularize_reflection(g, c.reg, roughness, c.path_roughness);
    DeltaSampled ds = delta_reflection_sampled(m, rr, anisotropy, c.info);
    if (ds == SPECULAR_PEAK_NOT_SAMPLED)
    {
        pdf = 0.0f;
        return col(0.0f);
    }
    float HoL = clampf_(1.0e-8f, 1.0f, dot(lh, ll));
    Color F = f82_tint_fresnel(m.base_color, m.metallic_F82, m.metallic_F90, m.metallic_F90_falloff_exponent, HoL);
    // Thin-film Fresnel is lerped in with weight material.thin_film (0 for
    // every material this build accepts).
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

struct LobeWeights {
    float coat, sheen, metal1, metal2, specular, diffuse, glass, dtrans;
};

// Principled.h:1515-1554 (outside the object: opaque scenes never enter one).
PT_HD LobeWeights lobe_weights(const Material& m)
{
    const float outside = 1.0f;
    LobeWeights w;
    w.coat = m.coat * outside;
    w.sheen = m.sheen * outside;
    float metallic = m.metallic;
    w.metal1 = metallic * outside;
    w.metal2 = metallic * outside;
    w.metal1 = lerpf_(w.metal1, 0.0f, m.second_roughness_weight);
    w.metal2 = lerpf_(0.0f, w.metal2, m.second_roughness_weight);
    w.glass = (1.0f - metallic) * (1.0f - m.diffuse_transmission) * m.specular_transmission;
    w.dtrans = (1.0f - metallic) * m.diffuse_transmission;
    w.specular = (1.0f - metallic) * (1.0f - m.specular_transmission * (1.0f - m.diffuse_transmission)) * m.specular *
                 outside;
    w.diffuse = (1.0f - metallic) * (1.0f - m.specular_transmission) * (1.0f - m.diffuse_transmission) * outside;
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

// Principled.h:1620-1715, for materials whose coat, sheen, specular, glass and
// diffuse-transmission weights are zero (the lobes this build implements:
// two GGX metal lobes and the Lambertian diffuse lobe).
PT_HD Color principled_eval(const BsdfGlobals& g, const Material& m, const BsdfCtx& c, float& pdf)
{
    pdf = 0.0f;
    bool refracting = dot(c.sn, c.to_light) < 0.0f;
    f3 T, B;
    build_onb(c.sn, T, B);
    f3 lv = world_to_local(T, B, c.sn, c.view);
    f3 ll = world_to_local(T, B, c.sn, c.to_light);
    f3 TR, BR;
    build_rotated_onb(c.sn, TR, BR, m.anisotropy_rotation * PI_F);
    f3 lvr = world_to_local(TR, BR, c.sn, c.view);
    f3 llr = world_to_local(TR, BR, c.sn, c.to_light);
    f3 lhr = normalize(lvr + llr);
    LobeWeights w = lobe_weights(m);
    LobeWeights p = lobe_probas(w);
    Color layers = col(1.0f);
    Color final_color = col(0.0f);
    // metal lobes (Principled.h:1129-1157)
    float wm[2] = {w.metal1 * (float)!refracting, w.metal2 * (float)!refracting};
    float pm[2] = {p.metal1, p.metal2};
    float rm[2] = {m.roughness, m.second_roughness};
    for (int k = 0; k < 2; k++)
    {
        if (wm[k] > 0.0f && lvr.z > 0.0f && llr.z > 0.0f)
        {
            float mpdf = 0.0f;
            Color contrib = metallic_eval(g, m, c, rm[k], m.anisotropy, lvr, llr, lhr, mpdf);
            contrib = contrib * wm[k];
            contrib = contrib * layers;
            pdf += mpdf * pm[k];
            final_color = final_color + contrib;
        }
        else
            final_color = final_color + col(0.0f);
    }
    // glossy base = specular (weight 0 here) + Lambertian diffuse
    // (Principled.h:1436-1492, Lambertian.h:14-23).
    Color glossy = col(0.0f) + col(0.0f);
    float wd = w.diffuse * (float)!refracting;
    if (wd > 0.0f && lv.z > 0.0f && ll.z > 0.0f)
    {
        float dpdf = 0.0f;
        Color contrib = col(0.0f);
        if (ll.z > 0.0f)
        {
            dpdf = ll.z * INV_PI_F;
            contrib = m.base_color * INV_PI_F;
        }
        contrib = contrib * wd;
        contrib = contrib * layers;
        pdf += dpdf * p.diffuse;
        glossy = glossy + contrib;
    }
    // glossy-base energy compensation is 1 without a specular layer
    // (PrincipledEnergyCompensation.h:21-25); clearcoat compensation is 1
    // without a coat (:84-88).
    final_color = final_color + glossy / 1.0f;
    return final_color / 1.0f;
}

PT_HD float principled_pdf(const BsdfGlobals& g, const Material& m, const BsdfCtx& c)
{
    float pdf = 0.0f;
    bool refracting = dot(c.sn, c.to_light) < 0.0f;
    f3 T, B;
    build_onb(c.sn, T, B);
    f3 lv = world_to_local(T, B, c.sn, c.view);
    f3 ll = world_to_local(T, B, c.sn, c.to_light);
    f3 TR, BR;
    build_rotated_onb(c.sn, TR, BR, m.anisotropy_rotation * PI_F);
    f3 lvr = world_to_local(TR, BR, c.sn, c.view);
    f3 llr = world_to_local(TR, BR, c.sn, c.to_light);
    f3 lhr = normalize(lvr + llr);
    LobeWeights w = lobe_weights(m);
    LobeWeights p = lobe_probas(w);
    float wm[2] = {w.metal1 * (float)!refracting, w.metal2 * (float)!refracting};
    float pm[2] = {p.metal1, p.metal2};
    float rm[2] = {m.roughness, m.second_roughness};
    for (int k = 0; k < 2; k++)
        if (wm[k] > 0.0f && lvr.z > 0.0f && llr.z > 0.0f)
            pdf += metallic_pdf(g, m, c, rm[k], m.anisotropy, lvr, llr, lhr) * pm[k];
    float wd = w.diffuse * (float)!refracting;
    if (wd > 0.0f && lv.z > 0.0f && ll.z > 0.0f)
        pdf += (ll.z > 0.0f ? ll.z * INV_PI_F : 0.0f) * p.diffuse;
    return pdf;
}

// Principled.h:1802-1939. Returns the BSDF value of the sampled direction;
// pdf = 0 when the sample is invalid. `c.info` receives the sampled lobe.
PT_HD Color principled_sample(const BsdfGlobals& g, const Material& m, BsdfCtx& c, f3& out_dir, float& pdf,
                              Xorshift32& rng)
{
    pdf = 0.0f;
    LobeWeights w = lobe_weights(m);
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
    f3 TR, BR;
    build_rotated_onb(c.sn, TR, BR, m.anisotropy_rotation * PI_F);
    f3 lvr = world_to_local(TR, BR, c.sn, c.view);
    // Coat, sheen, specular, diffuse-transmission and glass lobes are not
    // built; materials with those lobes are rejected at upload, so their
    // branches carry zero probability.
    if (rand_1 < cdf0)
        return col(0.0f);
    else if (rand_1 < cdf1)
        return col(0.0f);
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
        return col(0.0f);
    else if (rand_1 < cdf5)
    {
        c.info = FROM_DIFFUSE;
        out_dir = cosine_sample_around_normal(c.sn, rng);
    }
    else
        return col(0.0f);
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
