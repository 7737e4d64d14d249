/*
 * pathtrace.c -- CPU restatement of the reference's CameraRays + MegaKernel
 * for scenes of Lambertian and GGX-metal Principled materials (checker only).
 *
 * Test infrastructure: loaded only by tests/ and smoke(). It follows the
 * reference CPU renderer's per-pixel structure (CPURenderer.cpp:612-705 ->
 * CameraRays(render_data, x, y) then MegaKernel(render_data, x, y)), one pixel
 * at a time with OpenMP over rows, tracing with the octree restatement. Each
 * function names the reference file:line it restates.
 */
#include "oracle.h"
#include "fmath.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { float x, y, z; } v3;
typedef struct { float r, g, b; } rgb;

static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vscale(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vdivs(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float vlen(v3 a) { return sqrtf(vdot(a, a)); }
static inline v3 vnorm(v3 a) { float l = sqrtf(vdot(a, a)); return V(a.x / l, a.y / l, a.z / l); }
static inline v3 vneg(v3 a) { return V(-a.x, -a.y, -a.z); }

static inline rgb C3(float r, float g, float b) { rgb c = {r, g, b}; return c; }
static inline rgb C1(float v) { return C3(v, v, v); }
static inline rgb cadd(rgb a, rgb b) { return C3(a.r + b.r, a.g + b.g, a.b + b.b); }
static inline rgb csub(rgb a, rgb b) { return C3(a.r - b.r, a.g - b.g, a.b - b.b); }
static inline rgb cmul(rgb a, rgb b) { return C3(a.r * b.r, a.g * b.g, a.b * b.b); }
static inline rgb cscale(rgb a, float k) { return C3(a.r * k, a.g * k, a.b * k); }
static inline rgb cdivs(rgb a, float k) { return C3(a.r / k, a.g / k, a.b / k); }
static inline rgb cdiv(rgb a, rgb b) { return C3(a.r / b.r, a.g / b.g, a.b / b.b); }

/* hippt::max/min/clamp/lerp (src/HostDeviceCommon/Math.h:373-407, :527) */
static inline float hmax(float a, float b) { return a > b ? a : b; }
static inline float hmin(float a, float b) { return a < b ? a : b; }
static inline float hclamp(float lo, float hi, float v) { return hmin(hi, hmax(lo, v)); }
static inline float hlerp(float a, float b, float t) { return (1.0f - t) * a + t * b; }
static inline float cmax(rgb c) { return hmax(c.r, hmax(c.g, c.b)); } /* Color.h:98 */
static inline int cblack(rgb c) { return !(c.r > 0.0f || c.g > 0.0f || c.b > 0.0f); } /* Color.h:95 */
static inline float lum(rgb c) { return 0.3086f * c.r + 0.6094f * c.g + 0.0820f * c.b; } /* Color.h:90 */
static inline rgb clerp(rgb a, rgb b, float t) { return cadd(cscale(a, 1.0f - t), cscale(b, t)); } /* Math.h lerp */

#define M_PI_F 3.14159265358979323846f
#define M_TWO_PI_F 6.28318530717958647693f
#define M_FOUR_PI_F 12.5663706143591729539f
#define M_INV_PI_F 0.31830988618379067154f

/* ------------------------------------------------------------------ */
/* Materials: pack_to_GPU + unpack (MaterialCPU.h:31-150, Packing.h:56-203) */
/* ------------------------------------------------------------------ */
typedef struct {
    rgb emission, base_color;
    float roughness, metallic;
    rgb F82, F90;
    float F90_exp, anisotropy, anisotropy_rotation, second_roughness_weight, second_roughness;
    float specular, coat, sheen, specular_transmission, diffuse_transmission, thin_film, ior;
    int do_metallic_ec, do_specular_ec, do_coat_ec, do_glass_ec, thin_walled;
    float specular_tint;
    rgb specular_color;
    float specular_darkening;
    rgb coat_absorption;
    float coat_thickness, coat_roughness, coat_roughening, coat_darkening, coat_anisotropy, coat_anisotropy_rotation,
        coat_ior;
    float sheen_roughness;
    rgb sheen_color;
    float absorption_at_distance;
    rgb absorption_color;
    float dispersion_scale, dispersion_abbe_number;
    int dielectric_priority;
    float tf_ior, tf_thickness, tf_kappa3, tf_hue, tf_base_ior; /* thin film; hue = 8-bit degrees / 360 */
    int tf_override;
} omat;

static unsigned char q8(float x) { return (unsigned char)(x * 255.0f); }
static float u8f(unsigned char u) { return (float)u * (1.0f / 255.0f); }
static rgb color24(const float* c)
{
    /* ColorRGB24bFloat0_1Packed::get_color: g is stored shifted by 8 and read
     * back as float(g << 8) * (1 / (255 << 8)). */
    unsigned r = q8(c[0]), g = q8(c[1]), b = q8(c[2]);
    return C3((float)r * (1.0f / 255.0f), (float)(g << 8) * (1.0f / (float)(255 << 8)), (float)b * (1.0f / 255.0f));
}

/* CPUMaterial::make_safe (MaterialCPU.h:156-212), applied by the scene
 * parser to every material (SceneParser.cpp:437) */
static void clamp_c(float* c, float lo, float hi)
{
    for (int k = 0; k < 3; k++)
        c[k] = hclamp(lo, hi, c[k]);
}

static orc_cpu_material make_safe(orc_cpu_material m)
{
    clamp_c(m.base_color, 0.0f, 1.0f);
    m.metallic = hclamp(0.0f, 1.0f, m.metallic);
    clamp_c(m.metallic_F82, 0.0f, 1.0f);
    clamp_c(m.metallic_F90, 0.0f, 1.0f);
    m.anisotropy = hclamp(0.0f, 1.0f, m.anisotropy);
    m.anisotropy_rotation = hclamp(0.0f, 1.0f, m.anisotropy_rotation);
    m.second_roughness_weight = hclamp(0.0f, 1.0f, m.second_roughness_weight);
    m.second_roughness = hclamp(0.0f, 1.0f, m.second_roughness);
    m.specular = hclamp(0.0f, 1.0f, m.specular);
    m.specular_tint = hclamp(0.0f, 1.0f, m.specular_tint);
    clamp_c(m.specular_color, 0.0f, 1.0f);
    m.specular_darkening = hclamp(0.0f, 1.0f, m.specular_darkening);
    m.coat = hclamp(0.0f, 1.0f, m.coat);
    clamp_c(m.coat_medium_absorption, 0.0f, 1.0f);
    m.coat_roughness = hclamp(1.0e-4f, 1.0f, m.coat_roughness);
    m.coat_roughening = hclamp(0.0f, 1.0f, m.coat_roughening);
    m.coat_darkening = hclamp(0.0f, 1.0f, m.coat_darkening);
    m.coat_anisotropy = hclamp(0.0f, 1.0f, m.coat_anisotropy);
    m.coat_anisotropy_rotation = hclamp(0.0f, 1.0f, m.coat_anisotropy_rotation);
    m.sheen = hclamp(0.0f, 1.0f, m.sheen);
    m.sheen_roughness = hclamp(1.0e-4f, 1.0f, m.sheen_roughness);
    clamp_c(m.sheen_color, 0.0f, 1.0f);
    m.specular_transmission = hclamp(0.0f, 1.0f, m.specular_transmission);
    m.absorption_at_distance = hmax(m.absorption_at_distance, 1.0e-4f);
    for (int k = 0; k < 3; k++)
        m.absorption_color[k] = hmax(m.absorption_color[k], 1.0f / 512.0f);
    clamp_c(m.absorption_color, 0.0f, 1.0f);
    m.dispersion_abbe_number = hmax(1.0e-5f, m.dispersion_abbe_number);
    m.dispersion_scale = hclamp(0.0f, 1.0f, m.dispersion_scale);
    m.thin_film = hclamp(0.0f, 1.0f, m.thin_film);
    m.thin_film_hue_shift_degrees = hclamp(0.0f, 360.0f, m.thin_film_hue_shift_degrees);
    m.thin_film_ior = hmax(1.0005f, m.thin_film_ior);
    m.dielectric_priority = m.dielectric_priority < 0 ? 0 : m.dielectric_priority > 15 ? 15 : m.dielectric_priority;
    for (int k = 0; k < 3; k++)
        m.emission[k] = hmax(0.0f, m.emission[k]);
    if (m.specular_transmission == 0.0f && m.diffuse_transmission == 0.0f)
        m.dielectric_priority = (1 << 15) - 1; /* stored in 8 bits: 255 */
    return m;
}

static omat unpack_material(const orc_cpu_material* in)
{
    const orc_cpu_material safe = make_safe(*in);
    const orc_cpu_material* m = &safe;
    omat o;
    o.emission = C3(m->emission[0] * m->emission_strength * m->global_emissive_factor,
                    m->emission[1] * m->emission_strength * m->global_emissive_factor,
                    m->emission[2] * m->emission_strength * m->global_emissive_factor);
    o.base_color = color24(m->base_color);
    o.roughness = u8f(q8(m->roughness));
    o.metallic = u8f(q8(m->metallic));
    o.F82 = color24(m->metallic_F82);
    o.F90 = color24(m->metallic_F90);
    o.F90_exp = m->metallic_F90_falloff_exponent;
    o.anisotropy = u8f(q8(m->anisotropy));
    o.anisotropy_rotation = u8f(q8(m->anisotropy_rotation));
    o.second_roughness_weight = u8f(q8(m->second_roughness_weight));
    o.second_roughness = u8f(q8(m->second_roughness));
    o.specular = u8f(q8(m->specular));
    o.coat = u8f(q8(m->coat));
    o.sheen = u8f(q8(m->sheen));
    o.specular_transmission = u8f(q8(m->specular_transmission));
    o.diffuse_transmission = u8f(q8(m->diffuse_transmission));
    o.thin_film = u8f(q8(m->thin_film));
    o.ior = m->ior;
    o.do_metallic_ec = m->do_metallic_energy_compensation != 0;
    o.do_specular_ec = m->do_specular_energy_compensation != 0;
    o.do_coat_ec = m->do_coat_energy_compensation != 0;
    o.do_glass_ec = m->do_glass_energy_compensation != 0;
    o.thin_walled = m->thin_walled != 0;
    o.specular_tint = u8f(q8(m->specular_tint));
    o.specular_color = color24(m->specular_color);
    o.specular_darkening = u8f(q8(m->specular_darkening));
    o.coat_absorption = color24(m->coat_medium_absorption);
    o.coat_thickness = m->coat_medium_thickness;
    o.coat_roughness = u8f(q8(m->coat_roughness));
    o.coat_roughening = u8f(q8(m->coat_roughening));
    o.coat_darkening = u8f(q8(m->coat_darkening));
    o.coat_anisotropy = u8f(q8(m->coat_anisotropy));
    o.coat_anisotropy_rotation = u8f(q8(m->coat_anisotropy_rotation));
    o.coat_ior = m->coat_ior;
    o.sheen_roughness = u8f(q8(m->sheen_roughness));
    o.sheen_color = color24(m->sheen_color);
    o.absorption_at_distance = m->absorption_at_distance;
    o.absorption_color = color24(m->absorption_color);
    o.dispersion_scale = u8f(q8(m->dispersion_scale));
    o.dispersion_abbe_number = m->dispersion_abbe_number;
    o.dielectric_priority = (unsigned char)m->dielectric_priority;
    o.tf_ior = m->thin_film_ior;
    o.tf_thickness = m->thin_film_thickness;
    o.tf_kappa3 = m->thin_film_kappa_3;
    o.tf_hue = u8f(q8(m->thin_film_hue_shift_degrees / 360.0f));
    o.tf_base_ior = m->thin_film_base_ior_override;
    o.tf_override = m->thin_film_do_ior_override != 0;
    /* get_intersection_material (Material.h:122-149): a coat roughens the
     * base and second metal roughness */
    if (o.coat > 0.0f && o.coat_roughening > 0.0f)
    {
        float r4 = o.roughness * o.roughness, c4 = o.coat_roughness * o.coat_roughness;
        r4 = r4 * r4;
        c4 = c4 * c4;
        float target = sqrtf(sqrtf(hmin(1.0f, r4 + 2.0f * c4)));
        float rough = hlerp(o.roughness, target, o.coat);
        o.roughness = hlerp(o.roughness, rough, o.coat_roughening);
        if (o.second_roughness_weight > 0.0f)
        {
            float s4 = o.second_roughness * o.second_roughness;
            s4 = s4 * s4;
            float target2 = sqrtf(sqrtf(hmin(1.0f, s4 + 2.0f * c4)));
            float rough2 = hlerp(o.second_roughness, target2, o.coat);
            o.second_roughness = hlerp(o.second_roughness, rough2, o.coat_roughening);
        }
    }
    return o;
}

/* ------------------------------------------------------------------ */
/* Render context                                                       */
/* ------------------------------------------------------------------ */
typedef struct {
    const float* verts;
    const int32_t* idx;
    const int32_t* mat_idx;
    const omat* mats;
    const int32_t* emissive;
    int32_t ne;
    const float* areas;
    const uint8_t* has_normals; /* smooth shading; NULL: geometric normals */
    const float* vnormals;
    const float* lut;        /* GGX conductor E_ss, 128 x 128 */
    const float* lut_glossy; /* glossy dielectric E_ss, [128][64][128] */
    const float* lut_sheen;  /* sheen LTC (Ai, Bi, Ri), 32 x 32 */
    const float* lut_glass;      /* GGX glass E, [128][16][256], entering */
    const float* lut_glass_inv;  /* same, leaving */
    const float* lut_thin_glass; /* thin-walled glass E, [96][32][32] */
    const float* glass_exp;      /* correction exponent, [2][10][11][4] */
    const orc_octree* tree;
    const orc_camera* cam;
    orc_render_settings s;
    uint32_t sample_number;
    float consistent_tau;
    float consistent_tau_refract; /* consistent_tau(tau_0, sample_number + 1) */
    /* LSS_BASE_POWER: alias table over the emissive list */
    const float* alias_probas;
    const int32_t* alias_alias;
    float alias_sum;
    /* LSS_RESTIR_DI: reservoirs shaded at bounce 0 (struct ores, restir part) */
    const void* restir_out;
    int32_t final_shading_visibility;
    /* ReSTIR DI's later bounces by RIS (ReSTIR_DI_LaterBouncesSamplingStrategy) */
    int32_t ris_later;
    /* ambient light (WorldSettings.h:21-53): packed texels, alias table */
    int32_t env_type;
    float env_uniform[3];
    int32_t env_scale_bg;
    float env_intensity, env_sum;
    int32_t env_w, env_h;
    const uint32_t* env_texels;
    const float* env_probas;
    const int32_t* env_alias;
    float env_to_world[3][3], env_to_env[3][3];
} rctx;

enum { NO_INFO = 0, FROM_COAT = 2, FROM_FIRST_METAL = 4, FROM_SECOND_METAL = 8, FROM_SPECULAR = 16,
       FROM_GLASS_REFLECT = 32, FROM_GLASS_REFRACT = 64, FROM_DIFFUSE = 128, FROM_DIFFUSE_TRANSMISSION = 256 };
enum { NO_REG = 0, REG_CLASSIC = 1, REG_MIS = 2 };

static void tri(const rctx* R, int32_t p, v3* a, v3* b, v3* c)
{
    const int32_t* t = R->idx + 3 * (int64_t)p;
    const float* A = R->verts + 3 * (int64_t)t[0];
    const float* B = R->verts + 3 * (int64_t)t[1];
    const float* Cc = R->verts + 3 * (int64_t)t[2];
    *a = V(A[0], A[1], A[2]);
    *b = V(B[0], B[1], B[2]);
    *c = V(Cc[0], Cc[1], Cc[2]);
}

/* One closest-hit query through the octree, with an exclusive tmax. */
static orc_hit trace1(const rctx* R, v3 o, v3 d, int32_t last, float tmax)
{
    orc_ray ray;
    ray.o[0] = o.x; ray.o[1] = o.y; ray.o[2] = o.z;
    ray.d[0] = d.x; ray.d[1] = d.y; ray.d[2] = d.z;
    ray.last_prim = last;
    ray.tmax = tmax;
    orc_hit h;
    orc_octree_trace(R->tree, &ray, &h, 1, 0, 1);
    return h;
}

/* ------------------------------------------------------------------ */
/* ONB, frames, sampling (ONB.h:18-76, Sampling.h:147-201)             */
/* ------------------------------------------------------------------ */
static void build_onb(v3 N, v3* T, v3* B)
{
    if (N.z < -0.99998796f)
    {
        *T = V(0.0f, -1.0f, 0.0f);
        *B = V(-1.0f, 0.0f, 0.0f);
        return;
    }
    float nxa = -N.x / (1.0f + N.z);
    *T = V(1.0f + N.x * nxa, nxa * N.y, -N.x);
    *B = V(T->y, 1.0f - N.y * N.y / (1.0f + N.z), -N.y);
}

static void build_rotated_onb(v3 N, v3* T, v3* B, float rot)
{
    v3 up = fabsf(N.z) < 0.9999999f ? V(0.0f, 0.0f, 1.0f) : V(1.0f, 0.0f, 0.0f);
    v3 t = vnorm(vcross(up, N));
    float c = orc_cosf(rot), s = orc_sinf(rot);
    t = vadd(vadd(vscale(t, c), vscale(vcross(N, t), s)), vscale(vscale(N, vdot(N, t)), 1.0f - c));
    *T = t;
    *B = vcross(N, t);
}

static v3 to_local(v3 T, v3 B, v3 N, v3 v) { return vnorm(V(vdot(v, T), vdot(v, B), vdot(v, N))); }
static v3 to_world(v3 T, v3 B, v3 N, v3 v) { return vnorm(vadd(vadd(vscale(T, v.x), vscale(B, v.y)), vscale(N, v.z))); }
static v3 reflect(v3 d, v3 n) { return vsub(vscale(n, 2.0f * vdot(d, n)), d); }

static v3 cosine_around_normal(v3 n, uint32_t* st)
{
    float r1 = orc_rand01(st);
    float r2 = 2.0f * orc_rand01(st) - 1.0f;
    if (r1 < 1.0e-8f && r2 < -0.999999f && n.z > 0.999999f)
    {
        r1 += 1.0e-7f;
        r2 += 1.0e-7f;
    }
    float theta = M_TWO_PI_F * r1;
    float s = sqrtf(1.0f - r2 * r2);
    return vnorm(vadd(n, V(s * orc_cosf(theta), s * orc_sinf(theta), r2)));
}

/* ------------------------------------------------------------------ */
/* GGX (Microfacet.h), Fresnel F82 (Fresnel.h:138-168), energy           */
/* compensation (MicrofacetEnergyCompensation.h:27-45), regularisation   */
/* (MicrofacetRegularization.h:22-45)                                    */
/* ------------------------------------------------------------------ */
static void alphas(float r, float aniso, float* ax, float* ay)
{
    float aspect = sqrtf(1.0f - 0.9f * aniso);
    *ax = hmax(1.0e-4f, r * r / aspect);
    *ay = hmax(1.0e-4f, r * r * aspect);
}

static int smooth(float r) { return r <= 1.0e-2f; }

/* 0 not specular, 1 peak sampled, 2 peak not sampled (MaterialUtils.h:143-175) */
static int delta_sampled(const omat* m, float r, float aniso, int info)
{
    if (!smooth(r))
        return 0;
    int matching = fabsf(aniso - m->anisotropy) < 1.0e-3f;
    if (info == FROM_GLASS_REFLECT && smooth(m->roughness) && matching)
        return 1;
    int first = info == FROM_FIRST_METAL && smooth(m->roughness) && matching;
    int second = info == FROM_SECOND_METAL && smooth(m->second_roughness) && matching;
    if (first || second)
        return 1;
    if (info == FROM_COAT && fabsf(aniso - m->coat_anisotropy) < 1.0e-3f && smooth(m->coat_roughness))
        return 1;
    if (info == FROM_SPECULAR && smooth(m->roughness) && matching)
        return 1;
    return 2;
}

static float regularize(const rctx* R, int mode, float r, float path_r)
{
    if (!R->s.do_microfacet_regularization || mode == NO_REG)
        return r;
    float pt = R->consistent_tau / hmax(path_r * path_r, 1.0e-8f);
    float reg = sqrtf(sqrtf(1.0f / (pt * M_PI_F)));
    return hmax(R->s.microfacet_regularization_min_roughness, hmax(r, reg));
}

static float ggx_D(float ax, float ay, v3 h)
{
    float den = (h.x * h.x) / (ax * ax) + (h.y * h.y) / (ay * ay) + (h.z * h.z);
    return 1.0f / (M_PI_F * ax * ay * den * den);
}
static float lambda_(float ax, float ay, v3 d)
{
    float a = d.x * ax, b = d.y * ay;
    return (-1.0f + sqrtf(1.0f + (a * a + b * b) / (d.z * d.z))) * 0.5f;
}
static float vndf(float D, float G1V, v3 v, v3 h)
{
    float hol = hmax(1.0e-3f, vdot(v, h));
    return G1V * D * hol / v.z;
}

static float pow6(float x) { float x2 = x * x, x4 = x2 * x2; return x4 * x2; }

static rgb f82(rgb F0, rgb F82c, rgb F90, float e, float c)
{
    rgb base = cadd(F0, cscale(csub(F90, F0), orc_powf(1.0f - c, e)));
    if (cmax(base) < 1.0e-8f)
        return C1(0.0f);
    float laz = c * pow6(1.0f - c);
    const float cmx = 1.0f / 7.0f;
    const float denom_a = cmx * pow6(1.0f - cmx);
    rgb nume_a = cmul(cadd(F0, cscale(csub(F90, F0), orc_powf(1.0f - cmx, e))), csub(C1(1.0f), F82c));
    rgb a = cdivs(nume_a, denom_a);
    rgb F = csub(base, cscale(a, laz));
    return C3(hclamp(0.0f, 1.0f, F.r), hclamp(0.0f, 1.0f, F.g), hclamp(0.0f, 1.0f, F.b));
}

/* Nearest texel of the E_ss(cos, roughness) LUT; v = 1 - roughness is the
 * texture's own coordinate (row 0 = roughness 1). */
static float lut_ess(const float* lut, float u, float v)
{
    const int n = 128;
    if (u != 1.0f) u -= (float)(int)u;
    if (v != 1.0f) v -= (float)(int)v;
    u = u < 0 ? 1.0f + u : u;
    v = v < 0 ? 1.0f + v : v;
    int x = (int)(u * (float)(n - 1)), y = (int)(v * (float)(n - 1));
    return lut[x + y * n];
}

static rgb ec_conductor(const rctx* R, rgb F, float r, int do_ec, v3 lv)
{
    if (!do_ec || r <= 0.0f || lv.z < 0.0f)
        return C1(1.0f);
    float ess = lut_ess(R->lut, hmax(0.0f, lv.z), 1.0f - r);
    float kms = (1.0f - ess) / ess;
    return cadd(C1(1.0f), cscale(F, kms));
}

static rgb ggx_reflect_ss(float r, float aniso, rgb F, v3 lv, v3 ll, v3 lh, float* pdf, int ds)
{
    *pdf = 0.0f;
    if (smooth(r))
    {
        if (ds == 2)
            return C1(0.0f);
        if (vdot(reflect(lv, V(0, 0, 1)), ll) < 0.999999f)
        {
            *pdf = 0.0f;
            return C1(0.0f);
        }
        *pdf = 1.0e9f;
        return cdivs(cmul(C1(1.0e9f), F), fabsf(ll.z));
    }
    if (ll.z < 0.0f)
        return C1(0.0f);
    float ax, ay;
    alphas(r, aniso, &ax, &ay);
    float D = ggx_D(ax, ay, lh);
    float lv_ = lambda_(ax, ay, lv);
    float G1V = 1.0f / (1.0f + lv_);
    float Dv = vndf(D, G1V, lv, lh);
    float NoV = hmax(1.0e-3f, fabsf(lv.z)), NoL = hmax(1.0e-3f, fabsf(ll.z));
    *pdf = Dv / (4.0f * vdot(lv, lh));
    if (*pdf == 0.0f)
        return C1(0.0f);
    float G1L = 1.0f / (1.0f + lambda_(ax, ay, ll));
    float G2 = G1V * G1L;
    return cdivs(cscale(cscale(F, D), G2), 4.0f * NoL * NoV);
}

static v3 vndf_sample(v3 v, float ax, float ay, uint32_t* st)
{
    float r1 = orc_rand01(st);
    float r2 = orc_rand01(st);
    v3 Vh = vnorm(V(ax * v.x, ay * v.y, v.z));
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    v3 T1 = lensq > 0.0f ? vdivs(V(-Vh.y, Vh.x, 0.0f), sqrtf(lensq)) : V(1.0f, 0.0f, 0.0f);
    v3 T2 = vcross(Vh, T1);
    float r = sqrtf(r1), phi = M_TWO_PI_F * r2;
    float t1 = r * orc_cosf(phi), t2 = r * orc_sinf(phi);
    float s = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - s) * sqrtf(1.0f - t1 * t1) + s * t2;
    v3 Nh = vadd(vadd(vscale(T1, t1), vscale(T2, t2)), vscale(Vh, sqrtf(hmax(0.0f, 1.0f - t1 * t1 - t2 * t2))));
    return vnorm(V(ax * Nh.x, ay * Nh.y, hmax(0.0f, Nh.z)));
}

static v3 ggx_sample(float r, float aniso, v3 lv, uint32_t* st)
{
    float below = lv.z < 0 ? -1.0f : 1.0f;
    float ax, ay;
    alphas(r, aniso, &ax, &ay);
    v3 m = (ax <= 1.0e-4f && ay <= 1.0e-4f) ? V(0, 0, 1) : vndf_sample(vscale(lv, below), ax, ay, st);
    return vnorm(reflect(lv, vscale(m, below)));
}

/* ------------------------------------------------------------------ */
/* Principled BSDF, metal + diffuse lobes (Principled.h:125-215,        */
/* :1129-1157, :1436-1492, :1515-1618, :1620-1715, :1802-1939)           */
/* ------------------------------------------------------------------ */
/* Nested-dielectrics interior stack and the medium state of a path
 * (NestedDielectrics.h:125-255, RayVolumeState.h:14-80): 4 entries, entry 0
 * is the air. Entry word: priority bits 0-3, topmost bit 4, odd parity bit 5,
 * material index from bit 6 (all ones: the air). */
#define VOL_N 4
#define VOL_AIR (0xffffffffu >> 6)
typedef struct {
    uint32_t e[VOL_N];
    int pos;
    int incident, outgoing, inside;
    float distance, wavelength;
} ovol;

static uint32_t ve_mat(uint32_t e) { return e >> 6; }
static uint32_t ve_make(uint32_t mat, uint32_t prio, int odd, int top)
{
    return ((mat & VOL_AIR) << 6) | (odd ? 32u : 0u) | (top ? 16u : 0u) | (prio & 15u);
}

static void vol_init(ovol* v)
{
    for (int i = 0; i < VOL_N; i++)
        v->e[i] = ve_make(VOL_AIR, 0, 1, 1);
    v->pos = 0;
    v->incident = v->outgoing = -1;
    v->inside = 0;
    v->distance = 0.0f;
    v->wavelength = 0.0f;
}

/* returns 1 when the boundary is skipped (inside a medium of higher priority) */
static int vol_push(ovol* v, int mat, int prio)
{
    if (v->pos == VOL_N - 1)
        return 0;
    uint32_t m = (uint32_t)mat & VOL_AIR;
    int last;
    for (last = v->pos; last >= 0; last--)
        if (ve_mat(v->e[last]) != m && (v->e[last] & 16u) && (v->e[last] & 32u))
            break;
    int odd = 1;
    for (int i = v->pos; i >= 0; i--)
        if (ve_mat(v->e[i]) == m)
        {
            v->e[i] &= ~16u;
            odd = !(v->e[i] & 32u);
            break;
        }
    v->inside = !odd;
    if (v->pos < VOL_N - 1)
        v->pos++;
    v->e[v->pos] = ve_make(m, (uint32_t)prio, odd, 1);
    if (prio < (int)(v->e[last] & 15u))
        return 1;
    if (odd)
    {
        v->incident = (int)ve_mat(v->e[last]);
        v->outgoing = (int)m;
    }
    else
    {
        v->incident = (int)m;
        v->outgoing = (int)ve_mat(v->e[last]);
    }
    return 0;
}

static void vol_pop(ovol* v, int inside)
{
    uint32_t top = ve_mat(v->e[v->pos]);
    if (v->pos > 0)
        v->pos--;
    if (inside)
    {
        int prev;
        for (prev = v->pos; prev >= 0; prev--)
            if (ve_mat(v->e[prev]) == top)
                break;
        if (prev >= 0)
            for (int i = prev + 1; i <= v->pos; i++)
                v->e[i - 1] = v->e[i];
        if (v->pos > 0)
            v->pos--;
    }
    for (int i = v->pos; i >= 0; i--)
        if (ve_mat(v->e[i]) == top)
        {
            v->e[i] |= 16u;
            break;
        }
}

typedef struct { v3 view, sn, gn, L; int info, reg; float path_r; ovol* vs; int update; } bctx;

/* ------------------------------------------------------------------ */
/* Thin-film interference (ThinFilm.h:13-226; Belcour & Barla 2017)      */
/* ------------------------------------------------------------------ */
static rgb tf_sensitivity(float opd, float shift)
{
    float phase = 2.0f * M_PI_F * opd * 1.0e-6f;
    static const float val[3] = {5.4856e-13f, 4.4201e-13f, 5.2481e-13f};
    static const float pos[3] = {1.6810e+06f, 1.7953e+06f, 2.2084e+06f};
    static const float var[3] = {4.3278e+09f, 9.3046e+09f, 6.6121e+09f};
    float x[3];
    for (int k = 0; k < 3; k++)
        x[k] = val[k] * sqrtf(2.0f * M_PI_F * var[k]) * orc_cosf(pos[k] * phase + shift) *
               orc_expf(-1.0f * var[k] * phase * phase);
    x[0] += 9.7470e-14f * sqrtf(2.0f * M_PI_F * 4.5282e+09f) * orc_cosf(2.2399e+06f * phase + shift) *
            orc_expf(-4.5282e+09f * phase * phase);
    return C3(x[0] / 1.0685e-7f, x[1] / 1.0685e-7f, x[2] / 1.0685e-7f);
}

static void tf_phase(float ci, float e1, float e2, float k2, float* par, float* perp)
{
    float s2 = 1.0f - ci * ci;
    float A = (e2 * e2) * (1.0f - k2 * k2) - (e1 * e1) * s2;
    float t = 2.0f * (e2 * e2) * k2;
    float B = sqrtf(A * A + t * t);
    float U = (float)sqrt((double)(A + B) * 0.5);
    float V = sqrtf((B - A) * 0.5f);
    float ec = e1 * ci;
    *perp = orc_atan2f(2.0f * e1 * V * ci, U * U + V * V - ec * ec);
    float px = (e2 * e2) * (1.0f + k2 * k2) * ci;
    *par = orc_atan2f(2.0f * e1 * (e2 * e2) * ci * (2.0f * k2 * U - (1.0f - k2 * k2) * V),
                      px * px - (e1 * e1) * (U * U + V * V));
}

static void tf_conductor(float ci, float eta, float k, float* Rp, float* Rs)
{
    float c2 = ci * ci, s2 = 1.0f - c2;
    float t1 = eta * eta - k * k - s2;
    float ab = sqrtf(t1 * t1 + 4.0f * k * k * eta * eta);
    float a = sqrtf(0.5f * (ab + t1));
    float u1 = ab + c2, u2 = 2.0f * a * ci;
    *Rs = hclamp(0.0f, 1.0f, (u1 - u2) / (u1 + u2));
    float u3 = ab * c2 + s2 * s2, u4 = u2 * s2;
    *Rp = hclamp(0.0f, 1.0f, *Rs * (u3 - u4) / (u3 + u4));
}

static rgb hue_shift(rgb c, float deg)
{
    if (deg == 0.0f)
        return c;
    float ca = orc_cosf(deg / 180.0f * M_PI_F), sa = orc_sinf(deg / 180.0f * M_PI_F);
    const float s13 = 0.57735026918962576451f, o3 = 1.0f / 3.0f;
    float a00 = ca + (1.0f - ca) / 3.0f, a01 = o3 * (1.0f - ca) - s13 * sa, a02 = o3 * (1.0f - ca) + s13 * sa;
    float a10 = o3 * (1.0f - ca) + s13 * sa, a11 = ca + o3 * (1.0f - ca), a12 = o3 * (1.0f - ca) - s13 * sa;
    float a20 = o3 * (1.0f - ca) - s13 * sa, a21 = o3 * (1.0f - ca) + s13 * sa, a22 = ca + o3 * (1.0f - ca);
    return C3(hclamp(0.0f, 1.0f, c.r * a00 + c.g * a01 + c.b * a02), hclamp(0.0f, 1.0f, c.r * a10 + c.g * a11 + c.b * a12),
              hclamp(0.0f, 1.0f, c.r * a20 + c.g * a21 + c.b * a22));
}

static rgb thin_film(const omat* m, float amb, float HoL)
{
    if (m->thin_film == 0.0f)
        return C1(0.0f);
    float e1 = amb, e2 = m->tf_ior, e3 = m->tf_override ? m->tf_base_ior : m->ior;
    float k3 = m->tf_override ? m->tf_kappa3 : 0.0f;
    float R12p = 0, R12s = 0, T121p = 0, T121s = 0, R23p = 0, R23s = 0, c2 = 0.0f;
    float r = e1 / e2;
    float ct2 = 1.0f - (1.0f - HoL * HoL) * (r * r);
    if (ct2 <= 0.0f)
    {
        R12s = R12p = 1.0f;
    }
    else
    {
        c2 = sqrtf(ct2);
        tf_conductor(HoL, e2 / e1, 0.0f, &R12p, &R12s);
        tf_conductor(c2, e3 / e2, k3, &R23p, &R23s);
        T121p = 1.0f - R12p;
        T121s = 1.0f - R12s;
    }
    float D = m->tf_thickness / 1000.0f * c2;
    float p21p, p21s, p23p, p23s;
    tf_phase(HoL, e1, e2, 0.0f, &p21p, &p21s);
    tf_phase(c2, e2, e3, k3, &p23p, &p23s);
    p21p = M_PI_F - p21p;
    p21s = M_PI_F - p21s;
    float rp = sqrtf(R12p * R23p), rs = sqrtf(R12s * R23s);
    float Rsum = (T121p * T121p * R23p) / (1.0f - R12p * R23p);
    rgb I = C1(R12p + Rsum);
    float Cm = Rsum - T121p;
    for (int k = 1; k <= 2; k++)
    {
        Cm *= rp;
        I = cadd(I, cscale(cscale(tf_sensitivity((float)k * D, (float)k * (p23p + p21p)), 2.0f), Cm));
    }
    Rsum = (T121s * T121s * R23s) / (1.0f - R12s * R23s);
    I = cadd(I, C1(R12s + Rsum));
    Cm = Rsum - T121s;
    for (int k = 1; k <= 2; k++)
    {
        Cm *= rs;
        I = cadd(I, cscale(cscale(tf_sensitivity((float)k * D, (float)k * (p23s + p21s)), 2.0f), Cm));
    }
    I = cscale(I, 0.5f);
    float rr = 2.3646381f * I.r - 0.8965361f * I.g - 0.4680737f * I.b;
    float gg = -0.5151664f * I.r + 1.4264000f * I.g + 0.0887608f * I.b;
    float bb = 0.0052037f * I.r - 0.0144081f * I.g + 1.0092106f * I.b;
    I = C3(hclamp(0.0f, 1.0f, rr), hclamp(0.0f, 1.0f, gg), hclamp(0.0f, 1.0f, bb));
    return hue_shift(I, m->tf_hue * 360.0f);
}

static rgb metallic_eval(const rctx* R, const omat* m, const bctx* c, float r, float inc, v3 lv, v3 ll, v3 lh,
                         float* pdf)
{
    float rr = regularize(R, c->reg, r, c->path_r);
    int ds = delta_sampled(m, rr, m->anisotropy, c->info);
    if (ds == 2)
    {
        *pdf = 0.0f;
        return C1(0.0f);
    }
    float HoL = hclamp(1.0e-8f, 1.0f, vdot(lh, ll));
    rgb F = clerp(f82(m->base_color, m->F82, m->F90, m->F90_exp, HoL), thin_film(m, inc, HoL), m->thin_film);
    rgb ms = ec_conductor(R, F, rr, m->do_metallic_ec, lv);
    rgb ss = ggx_reflect_ss(rr, m->anisotropy, F, lv, ll, lh, pdf, ds);
    return cmul(ss, ms);
}

typedef struct { float coat, sheen, m1, m2, spec, diff, glass, dt; } lobes;

static lobes weights(const omat* m, int outside_object)
{
    lobes w;
    const float outside = outside_object ? 1.0f : 0.0f;
    w.coat = m->coat * outside;
    w.sheen = m->sheen * outside;
    w.m1 = m->metallic * outside;
    w.m2 = m->metallic * outside;
    w.m1 = hlerp(w.m1, 0.0f, m->second_roughness_weight);
    w.m2 = hlerp(0.0f, w.m2, m->second_roughness_weight);
    w.glass = !outside_object ? (1.0f - m->diffuse_transmission)
                              : (1.0f - m->metallic) * (1.0f - m->diffuse_transmission) * m->specular_transmission;
    w.dt = !outside_object ? m->diffuse_transmission : (1.0f - m->metallic) * m->diffuse_transmission;
    w.spec = (1.0f - m->metallic) * (1.0f - m->specular_transmission * (1.0f - m->diffuse_transmission)) * m->specular * outside;
    w.diff = (1.0f - m->metallic) * (1.0f - m->specular_transmission) * (1.0f - m->diffuse_transmission) * outside;
    return w;
}

static lobes probas(lobes w)
{
    float nf = 1.0f / (w.coat + w.sheen + w.m1 + w.m2 + w.spec + w.diff + w.glass + w.dt);
    lobes p = {w.coat * nf, w.sheen * nf, w.m1 * nf, w.m2 * nf, w.spec * nf, w.diff * nf, w.glass * nf, w.dt * nf};
    return p;
}

/* Dielectric Fresnel (Fresnel.h:11-62, :169-172) */
static float f0_eta(float eta_t, float eta_i)
{
    float n = eta_t - eta_i, d = eta_t + eta_i;
    return (n * n) / (d * d);
}
static float fresnel_d(float cos_i, float rel)
{
    if (fabsf(1.0f - rel) < 1.0e-4f)
        return 0.0f;
    float si2 = 1.0f - cos_i * cos_i;
    float st2 = si2 / (rel * rel);
    if (st2 >= 1.0f)
        return 1.0f;
    float ct = sqrtf(1.0f - st2);
    float rpa = (rel * cos_i - ct) / (rel * cos_i + ct);
    float rpe = (cos_i - rel * ct) / (cos_i + rel * ct);
    return (rpa * rpa + rpe * rpe) / 2.0f;
}
static float fresnel_d2(float cos_i, float eta_i, float eta_t) { return fresnel_d(cos_i, eta_t / eta_i); }
static float hemi_albedo_fit(float rel)
{
    return orc_logf((10893.0f * rel - 1438.2f) / (-774.4f * (rel * rel) + 10212.0f * rel + 1.0f));
}
static rgb fdivc(float a, rgb b) { return C3(a / b.r, a / b.g, a / b.b); }

/* Stacked 3D LUT through the CPU nearest sampler (Image.cpp:761-803) */
static float lut3d(const float* lut, int w, int h, int d, float u, float v, float t)
{
    if (u != 1.0f) u -= (float)(int)u;
    if (v != 1.0f) v -= (float)(int)v;
    if (t != 1.0f) t -= (float)(int)t;
    u = u < 0 ? 1.0f + u : u;
    v = v < 0 ? 1.0f + v : v;
    t = t < 0 ? 1.0f + t : t;
    v = 1.0f - v;
    int x = (int)(u * (float)(w - 1)), y = (int)(v * (float)(h - 1)), z = (int)(t * (float)(d - 1));
    return lut[(size_t)z * (size_t)(w * h) + (size_t)(x + y * w)];
}

/* Specular layer and glossy base (Principled.h:215-330, :1272-1512,
 * PrincipledEnergyCompensation.h:15-125) */
static float spec_rel_ior(const omat* m, float inc)
{
    if (m->coat == 0.0f)
        return m->ior;
    float layer = hlerp(inc, m->coat_ior, m->coat);
    float rel = m->ior / layer;
    return rel < 1.0f ? 1.0f / rel : rel;
}

/* principled_specular_fresnel (Principled.h:215-241) */
static rgb spec_fresnel(const omat* m, float rel, float cos_i)
{
    rgb Fs = m->thin_film < 1.0f ? C1(fresnel_d(cos_i, rel)) : C1(0.0f);
    return clerp(Fs, thin_film(m, m->ior / rel, cos_i), m->thin_film);
}

static float ec_glossy(const rctx* R, const omat* m, float inc, float NoV)
{
    if (!m->do_specular_ec || m->roughness < 0.0f || m->specular < 1.0e-3f || NoV < 0.0f)
        return 1.0f;
    float rel = spec_rel_ior(m, inc);
    if (fabsf(rel - 1.0f) < 1.0e-3f)
        rel += 1.0e-3f;
    float u = orc_powf(NoV, 1.0f / 2.5f);
    float w = sqrtf(sqrtf(f0_eta(m->ior, m->ior / rel)));
    float ms = lut3d(R->lut_glossy, 128, 64, 128, u, m->roughness, w);
    float comp = hlerp(1.0f, ms, m->specular);
    return hlerp(comp, 1.0f, m->thin_film);
}

static float ec_clearcoat(const rctx* R, const omat* m, float inc, float NoV)
{
    if (!m->do_specular_ec || m->coat < 1.0e-3f || NoV < 0.0f)
        return 1.0f;
    if (fabsf(m->coat_ior / inc - 1.0f) < 1.0e-3f)
        inc += 1.0e-3f;
    float u = orc_powf(NoV, 1.0f / 2.5f);
    float w = sqrtf(sqrtf(f0_eta(m->coat_ior, inc)));
    float ms = lut3d(R->lut_glossy, 128, 64, 128, u, m->coat_roughness, w);
    float comp = hlerp(1.0f, ms, m->coat * (1.0f - m->specular_transmission));
    return hlerp(comp, 1.0f, m->thin_film);
}

static rgb spec_darkening(const omat* m, float rel, float vf)
{
    (void)vf;
    if (m->specular_darkening == 0.0f)
        return C1(1.0f);
    float K = 1.0f - (1.0f - hemi_albedo_fit(rel)) / (rel * rel);
    rgb dk = fdivc(1.0f - K, csub(C1(1.0f), cscale(m->base_color, K)));
    return clerp(C1(1.0f), dk, m->specular * m->specular_darkening * (1.0f - m->diffuse_transmission));
}

static rgb coat_darkening(const omat* m, float rel, float vf)
{
    if (m->coat_darkening == 0.0f)
        return C1(1.0f);
    float Kr = 1.0f - (1.0f - hemi_albedo_fit(rel)) / (rel * rel);
    float base_r = hlerp(1.0f, m->roughness, hmax(m->specular_transmission, hmax(m->metallic, m->specular)));
    float K = hlerp(vf, Kr, base_r);
    rgb albedo = cdivs(cadd(m->base_color, cscale(m->sheen_color, m->sheen)), 1.0f + m->sheen);
    rgb dk = fdivc(1.0f - K, csub(C1(1.0f), cscale(albedo, K)));
    return clerp(C1(1.0f), dk, m->coat * m->coat_darkening * (1.0f - m->diffuse_transmission));
}

/* Sheen LTC (SheenLTC.h) */
static rgb ltc_params(const rctx* R, float roughness, float cos_t)
{
    float u = cos_t, v = hclamp(0.0f, 1.0f, roughness);
    if (u != 1.0f) u -= (float)(int)u;
    if (v != 1.0f) v -= (float)(int)v;
    u = u < 0 ? 1.0f + u : u;
    v = v < 0 ? 1.0f + v : v;
    v = 1.0f - v;
    int x = (int)(u * 31.0f), y = (int)(v * 31.0f);
    const float* p = R->lut_sheen + 3 * (x + y * 32);
    return C3(p[0], p[1], p[2]);
}
static float ltc_eval(v3 l, rgb A)
{
    v3 o = V(l.x * A.r + l.z * A.g, l.y * A.r, l.z);
    float len = sqrtf(vdot(o, o));
    o = vdivs(o, len);
    float jac = (A.r * A.r) / (len * len * len);
    return o.z * M_INV_PI_F * jac;
}
static float phi_of(v3 d)
{
    float p = orc_atan2f(d.y, d.x);
    return p < 0.0f ? p + M_TWO_PI_F : p;
}
static v3 rotate_v(v3 v, v3 axis, float angle)
{
    float s, c;
    orc_sincos(angle, &s, &c);
    return vadd(vadd(vscale(v, c), vscale(vscale(axis, vdot(v, axis)), 1.0f - c)), vscale(vcross(axis, v), s));
}
static rgb sheen_eval(const rctx* R, const omat* m, v3 lv, v3 ll, float* pdf, float* refl)
{
    if (lv.z <= 0.0f || ll.z <= 0.0f)
    {
        *pdf = 0.0f;
        *refl = lv.z > 0.0f ? ltc_params(R, m->sheen_roughness, lv.z).b : 0.0f;
        return C1(0.0f);
    }
    v3 ls = rotate_v(ll, V(0, 0, 1), -phi_of(lv));
    rgb A = ltc_params(R, m->sheen_roughness, lv.z);
    float Do = ltc_eval(ls, A);
    *pdf = Do;
    *refl = A.b;
    return cdivs(cscale(cscale(m->sheen_color, A.b), Do), ll.z);
}
static float sheen_pdf(const rctx* R, const omat* m, v3 lv, v3 ll)
{
    if (lv.z <= 0.0f || ll.z <= 0.0f)
        return 0.0f;
    v3 ls = rotate_v(ll, V(0, 0, 1), -phi_of(lv));
    return ltc_eval(ls, ltc_params(R, m->sheen_roughness, lv.z));
}
static v3 sheen_sample(const rctx* R, const omat* m, v3 lv, uint32_t* st)
{
    float r1 = orc_rand01(st), r2 = orc_rand01(st);
    float phi = M_TWO_PI_F * r1, ct = sqrtf(r2), stt = sqrtf(1.0f - ct * ct);
    float s, c;
    orc_sincos(phi, &s, &c);
    v3 cs = vnorm(V(c * stt, s * stt, ct));
    rgb A = ltc_params(R, m->sheen_roughness, lv.z);
    float ai = 1.0f / A.r, bi = A.g;
    v3 d = vnorm(V(cs.x * ai - cs.z * bi * ai, cs.y * ai, cs.z));
    return rotate_v(d, V(0, 0, 1), phi_of(lv));
}

typedef struct { v3 lv, ll, lh, lvr, llr, lhr; } frames;

static frames make_frames(const omat* m, const bctx* c)
{
    frames f;
    v3 T, B, TR, BR;
    build_onb(c->sn, &T, &B);
    f.lv = to_local(T, B, c->sn, c->view);
    f.ll = to_local(T, B, c->sn, c->L);
    f.lh = vnorm(vadd(f.lv, f.ll));
    build_rotated_onb(c->sn, &TR, &BR, m->anisotropy_rotation * M_PI_F);
    f.lvr = to_local(TR, BR, c->sn, c->view);
    f.llr = to_local(TR, BR, c->sn, c->L);
    f.lhr = vnorm(vadd(f.lvr, f.llr));
    return f;
}

/* one GGX reflection lobe of a dielectric interface (coat or specular) */
static rgb ggx_dielectric(const rctx* R, const omat* m, const bctx* c, float rough, float aniso, rgb F, v3 lv, v3 ll,
                          v3 lh, float* pdf)
{
    float rr = regularize(R, c->reg, rough, c->path_r);
    return ggx_reflect_ss(rr, aniso, F, lv, ll, lh, pdf, delta_sampled(m, rr, aniso, c->info));
}
static float ggx_dielectric_pdf(const rctx* R, const omat* m, const bctx* c, float rough, float aniso, v3 lv, v3 ll,
                                v3 lh)
{
    float rr = regularize(R, c->reg, rough, c->path_r);
    int ds = delta_sampled(m, rr, aniso, c->info);
    float pdf;
    if (smooth(rr))
    {
        if (ds == 2 || vdot(reflect(lv, V(0, 0, 1)), ll) < 0.999999f)
            return 0.0f;
        return 1.0e9f;
    }
    if (ll.z < 0.0f)
        return 0.0f;
    float ax, ay;
    alphas(rr, aniso, &ax, &ay);
    float D = ggx_D(ax, ay, lh);
    float G1V = 1.0f / (1.0f + lambda_(ax, ay, lv));
    pdf = vndf(D, G1V, lv, lh) / (4.0f * vdot(lv, lh));
    return pdf;
}

static float metal_pdf(const rctx* R, const omat* m, const bctx* c, float r, v3 lv, v3 ll, v3 lh)
{
    float rr = regularize(R, c->reg, r, c->path_r);
    int ds = delta_sampled(m, rr, m->anisotropy, c->info);
    if (ds == 2)
        return 0.0f;
    return ggx_dielectric_pdf(R, m, c, r, m->anisotropy, lv, ll, lh);
}

/* ------------------------------------------------------------------ */
/* Glass and diffuse transmission (Principled.h:332-873, Microfacet.h:   */
/* 288-392, MicrofacetEnergyCompensation.h:101-703, MicrofacetRegular-   */
/* ization.h:60-146, MaterialUtils.h:32-57, Dispersion.h:433-535)        */
/* ------------------------------------------------------------------ */
static v3 refract_dir(v3 d, v3 n, float eta)
{
    float c = vdot(d, n);
    float s2 = 1.0f - c * c;
    float ct = sqrtf(1.0f - s2 / (eta * eta));
    return vadd(vdivs(vneg(d), eta), vscale(n, c / eta - ct));
}

static float thin_rough(int thin, float r, float eta)
{
    if (!thin)
        return r;
    float a = (eta - 0.5f) * (eta - 0.5f), b = eta * eta * eta;
    return hclamp(0.0f, 1.0f, r * sqrtf(3.7f * (eta - 1.0f) * a / b) / 1.39f);
}

static float reg_refract(const rctx* R, int mode, float r, float path_r, float ei, float et)
{
    if (!R->s.do_microfacet_regularization || mode == NO_REG)
        return r;
    float tau = R->consistent_tau_refract / hmax(path_r * path_r, 1.0e-8f);
    float d = ei - et, mx = hmax(ei, et);
    float reg = sqrtf(sqrtf(1.0f / (tau * M_PI_F * (d * d) / (4.0f * (mx * mx)))));
    return hmax(R->s.microfacet_regularization_min_roughness, hmax(r, reg));
}

static float reg_mix(const rctx* R, int mode, float r, float path_r, float ei, float et)
{
    if (!R->s.do_microfacet_regularization || mode == NO_REG)
        return r;
    float tau = R->consistent_tau / hmax(path_r * path_r, 1.0e-8f);
    float refl = sqrtf(sqrtf(1.0f / (tau * M_PI_F)));
    if (ei == et)
        return refl;
    float d = ei - et, mx = hmax(ei, et);
    float refr = sqrtf(sqrtf(1.0f / (tau * M_PI_F * (d * d) / (4.0f * (mx * mx)))));
    return hmax(R->s.microfacet_regularization_min_roughness, hmax(r, refr * 0.75f + refl * 0.25f));
}

/* one IOR knot of the hand-tabulated correction exponent */
static float exp_knot(const float* seg, float r)
{
    for (int k = 0; k < 11; k++)
        if (r <= seg[4 * k])
            return seg[4 * k + 1] == 0.0f ? seg[4 * k + 2]
                                          : hlerp(seg[4 * k + 2], seg[4 * k + 3], (r - seg[4 * (k - 1)]) / 0.1f);
    return 2.5f;
}

static float correction_exponent(const float* tab, float r, float eta)
{
    static const float knot[10] = {1.01f, 1.02f, 1.03f, 1.1f, 1.2f, 1.4f, 1.5f, 2.0f, 2.4f, 3.0f};
    if ((r < 1.0e-10f && r > -1.0e-10f) || fabsf(1.0f - eta) < 1.0e-3f)
        return 2.5f;
    float lo_b = 1.01f, lo = 2.5f;
    for (int k = 0; k < 9; k++)
        if (eta > knot[k] && eta <= knot[k + 1])
        {
            lo_b = knot[k];
            lo = exp_knot(tab + 44 * k, r);
            break;
        }
    int h = 9;
    for (int k = 0; k < 9; k++)
        if (eta <= knot[k])
        {
            h = k;
            break;
        }
    float hi = exp_knot(tab + 44 * (10 + h), r);
    if (knot[h] == lo_b)
        return lo;
    return hlerp(lo, hi, (eta - lo_b) / (knot[h] - lo_b));
}

static float ec_glass(const rctx* R, const omat* m, float r, int inside, float et, float ei, float eta, float NoV)
{
    if (!m->do_glass_ec || r <= 0.0f)
        return 1.0f;
    float comp = 1.0f;
    if (m->thin_film < 1.0f)
    {
        float e_corr = inside ? 1.0f / eta : eta;
        float ex = m->thin_walled ? 2.5f : correction_exponent(R->glass_exp, r, e_corr);
        float u = orc_powf(hmax(1.0e-3f, NoV), 1.0f / ex);
        float f0 = sqrtf(sqrtf(f0_eta(et, ei)));
        if (m->thin_walled)
            comp = lut3d(R->lut_thin_glass, 32, 32, 96, u, r, f0);
        else
            comp = lut3d(inside ? R->lut_glass_inv : R->lut_glass, 256, 16, 128, u, r, f0);
        comp = hlerp(comp, 1.0f, m->thin_film * r);
    }
    return comp;
}

static rgb beer(const rctx* R, ovol* v)
{
    const omat* md = &R->mats[v->incident];
    rgb a = md->absorption_color;
    if (a.r == 1.0f && a.g == 1.0f && a.b == 1.0f)
        return C1(1.0f);
    v->distance = hmax(v->distance, 1.0e-6f);
    rgb k = cdivs(C3(orc_logf(a.r), orc_logf(a.g), orc_logf(a.b)), md->absorption_at_distance);
    rgb e = cscale(k, v->distance);
    return C3(orc_expf(e.r), orc_expf(e.g), orc_expf(e.b));
}

static rgb ggx_refract(const omat* m, float r, float eta, rgb F, v3 lv, v3 ll, v3 lh, float* pdf, int info)
{
    float NoL = ll.z, NoV = lv.z, HoL = vdot(ll, lh), HoV = vdot(lv, lh);
    rgb col;
    if (smooth(r))
    {
        int perfect = vdot(refract_dir(lv, V(0, 0, 1), eta), ll) > 0.999999f;
        if (info == FROM_GLASS_REFRACT && perfect)
        {
            col = cmul(cmul(C1(1.0e9f), csub(C1(1.0f), F)), m->base_color);
            col = cdivs(col, fabsf(NoL));
            *pdf = 1.0e9f;
        }
        else
        {
            col = C1(0.0f);
            *pdf = 0.0f;
        }
    }
    else
    {
        float dp = HoL + HoV / eta, dp2 = dp * dp;
        float denom = dp2 * NoL * NoV;
        float ax, ay;
        alphas(r, m->anisotropy, &ax, &ay);
        float D = ggx_D(ax, ay, lh);
        float g1v = 1.0f / (1.0f + lambda_(ax, ay, lv)), g1l = 1.0f / (1.0f + lambda_(ax, ay, ll));
        float G2 = g1v * g1l;
        *pdf = (fabsf(HoL) / dp2) * (g1v / fabsf(NoV) * D * fabsf(HoV));
        col = cscale(cscale(cmul(cscale(m->base_color, D), csub(C1(1.0f), F)), G2), fabsf(HoL * HoV / denom));
    }
    return cdivs(col, eta * eta);
}

static float ggx_refract_pdf(const omat* m, float r, float eta, v3 lv, v3 ll, v3 lh, int info)
{
    float NoV = lv.z, HoL = vdot(ll, lh), HoV = vdot(lv, lh);
    if (smooth(r))
        return info == FROM_GLASS_REFRACT && vdot(refract_dir(lv, V(0, 0, 1), eta), ll) > 0.999999f ? 1.0e9f : 0.0f;
    float dp = HoL + HoV / eta, dp2 = dp * dp;
    float ax, ay;
    alphas(r, m->anisotropy, &ax, &ay);
    float D = ggx_D(ax, ay, lh);
    float g1v = 1.0f / (1.0f + lambda_(ax, ay, lv));
    return (fabsf(HoL) / dp2) * (g1v / fabsf(NoV) * D * fabsf(HoV));
}

/* Cauchy dispersion fitted to the Abbe number (OpenPBR), Dispersion.h:485-501 */
static float disp_ior(float abbe, float scale, float ior, float w)
{
    if (scale == 0.0f)
        return ior;
    float ab = abbe / scale;
    float B = (ior - 1.0f) / (ab * 0.00000191038851931481f);
    float A = ior - B / 334777.96f;
    return A + B / (w * w);
}

static float medium_ior(const rctx* R, int medium) { return (uint32_t)medium == VOL_AIR ? 1.0f : R->mats[medium].ior; }

static void glass_etas(const rctx* R, const omat* m, const ovol* v, float* ei, float* et, float* eta)
{
    *ei = disp_ior(m->dispersion_abbe_number, m->dispersion_scale, medium_ior(R, v->incident), fabsf(v->wavelength));
    *et = disp_ior(m->dispersion_abbe_number, m->dispersion_scale, medium_ior(R, v->outgoing), fabsf(v->wavelength));
    *eta = *et / *ei;
    if (fabsf(*eta - 1.0f) < 1.0e-5f)
        *eta = 1.0f + 1.0e-5f;
}

static rgb glass_F(const omat* m, float r, float ei, float eta, float HoV, float* prob)
{
    rgb tf = thin_film(m, ei, HoV);
    rgb F = clerp(m->thin_film < 1.0f ? C1(fresnel_d(HoV, eta)) : C1(0.0f), tf, m->thin_film);
    if (m->thin_walled && F.r < 1.0f && m->thin_film == 0.0f && r < 0.1f)
        F = cadd(F, C1((1.0f - F.r) * (1.0f - F.r) * F.r / (1.0f - F.r * F.r)));
    *prob = lum(F);
    return F;
}

static v3 glass_h(const rctx* R, int refl, int thin, float r, float eta, v3 lv, v3 ll)
{
    v3 h;
    if (r <= 1.0e-4f && !R->s.do_microfacet_regularization)
        h = V(0, 0, 1);
    else
        h = vnorm(refl ? vadd(ll, lv) : thin ? vadd(V(ll.x, ll.y, -ll.z), lv) : vadd(vscale(ll, eta), lv));
    return h.z < 0.0f ? vneg(h) : h;
}

static rgb glass_eval(const rctx* R, const omat* m, const bctx* c, v3 lv, v3 ll, float* pdf)
{
    *pdf = 0.0f;
    if (fabsf(ll.z) < 1.0e-8f)
        return C1(0.0f);
    int refl = ll.z * lv.z > 0.0f;
    ovol* v = c->vs;
    float ei, et, eta;
    glass_etas(R, m, v, &ei, &et, &eta);
    float sr = thin_rough(m->thin_walled, m->roughness, eta);
    v3 h = glass_h(R, refl, m->thin_walled, sr, eta, lv, ll);
    float HoL = vdot(ll, h), HoV = vdot(lv, h);
    if (HoL * ll.z < 0.0f || HoV * lv.z < 0.0f)
        return C1(0.0f);
    float prob;
    rgb F = glass_F(m, sr, ei, eta, HoV, &prob);
    int reg = R->s.do_microfacet_regularization;
    if (refl)
    {
        float rr = reg && c->reg == REG_MIS ? reg_mix(R, c->reg, sr, c->path_r, ei, et)
                 : reg && c->reg == REG_CLASSIC ? regularize(R, c->reg, sr, c->path_r) : sr;
        rgb col = ggx_reflect_ss(rr, m->anisotropy, F, lv, ll, h, pdf, delta_sampled(m, sr, m->anisotropy, c->info));
        col = cdivs(col, ec_glass(R, m, m->roughness, v->inside, et, ei, eta, lv.z));
        *pdf *= prob;
        return col;
    }
    float rr = reg && c->reg == REG_MIS ? reg_mix(R, c->reg, sr, c->path_r, ei, et)
             : reg && c->reg == REG_CLASSIC ? reg_refract(R, c->reg, sr, c->path_r, ei, et) : sr;
    rgb col = ggx_refract(m, rr, eta, F, lv, ll, h, pdf, c->info);
    *pdf *= 1.0f - prob;
    col = cdivs(col, ec_glass(R, m, rr, v->inside, et, ei, eta, lv.z));
    if (m->thin_walled)
        col = cmul(col, m->base_color);
    if (m->thin_walled && c->update)
        vol_pop(v, v->inside);
    else if ((uint32_t)v->incident != VOL_AIR)
    {
        col = cmul(col, beer(R, v));
        if (c->update)
        {
            v->distance = 0.0f;
            if (v->inside)
                vol_pop(v, 1);
        }
    }
    return col;
}

static float glass_pdf(const rctx* R, const omat* m, const bctx* c, v3 lv, v3 ll)
{
    if (fabsf(ll.z) < 1.0e-8f)
        return 0.0f;
    int refl = ll.z * lv.z > 0.0f;
    float ei, et, eta;
    glass_etas(R, m, c->vs, &ei, &et, &eta);
    float sr = thin_rough(m->thin_walled, m->roughness, eta);
    v3 h = glass_h(R, refl, m->thin_walled, sr, eta, lv, ll);
    float HoL = vdot(ll, h), HoV = vdot(lv, h);
    if (HoL * ll.z < 0.0f || HoV * lv.z < 0.0f)
        return 0.0f;
    float prob;
    glass_F(m, sr, ei, eta, HoV, &prob);
    int reg = R->s.do_microfacet_regularization;
    if (refl)
    {
        float rr = reg && c->reg == REG_MIS ? reg_mix(R, c->reg, sr, c->path_r, ei, et)
                 : reg && c->reg == REG_CLASSIC ? regularize(R, c->reg, sr, c->path_r) : sr;
        int ds = delta_sampled(m, sr, m->anisotropy, c->info);
        float pdf;
        if (smooth(rr))
            pdf = (ds == 2 || vdot(reflect(lv, V(0, 0, 1)), ll) < 0.999999f) ? 0.0f : 1.0e9f;
        else if (ll.z < 0.0f)
            pdf = 0.0f;
        else
        {
            float ax, ay;
            alphas(rr, m->anisotropy, &ax, &ay);
            float D = ggx_D(ax, ay, h);
            float G1V = 1.0f / (1.0f + lambda_(ax, ay, lv));
            pdf = vndf(D, G1V, lv, h) / (4.0f * vdot(lv, h));
        }
        return pdf * prob;
    }
    float rr = reg && c->reg == REG_MIS ? reg_mix(R, c->reg, sr, c->path_r, ei, et)
             : reg && c->reg == REG_CLASSIC ? reg_refract(R, c->reg, sr, c->path_r, ei, et) : sr;
    return ggx_refract_pdf(m, rr, eta, lv, ll, h, c->info) * (1.0f - prob);
}

static v3 glass_sample(const rctx* R, const omat* m, bctx* c, v3 lv, uint32_t* st)
{
    ovol* v = c->vs;
    float ei, et, eta;
    glass_etas(R, m, v, &ei, &et, &eta);
    float sr = thin_rough(m->thin_walled, m->roughness, eta);
    if (R->s.do_microfacet_regularization && c->reg == REG_MIS)
        sr = reg_mix(R, c->reg, sr, c->path_r, ei, et);
    float ax, ay;
    alphas(sr, m->anisotropy, &ax, &ay);
    v3 mn = (ax <= 1.0e-4f && ay <= 1.0e-4f) ? V(0, 0, 1) : vndf_sample(lv, ax, ay, st);
    float prob;
    glass_F(m, sr, ei, eta, vdot(lv, mn), &prob);
    if (orc_rand01(st) < prob)
    {
        c->info = FROM_GLASS_REFLECT;
        if (c->update)
            vol_pop(v, 0);
        return reflect(lv, mn);
    }
    c->info = FROM_GLASS_REFRACT;
    if (vdot(mn, lv) < 0.0f)
        mn = vneg(mn);
    if (m->thin_walled)
    {
        v3 r = reflect(lv, mn);
        r.z = r.z * -1.0f;
        if (c->update)
            vol_pop(v, 0);
        return r;
    }
    return refract_dir(lv, mn, eta);
}

static rgb dt_eval(const rctx* R, const omat* m, const bctx* c, v3 lv, v3 ll, float* pdf)
{
    *pdf = 0.0f;
    if (lv.z * ll.z > 0.0f)
        return C1(0.0f);
    rgb col = cscale(m->base_color, M_INV_PI_F);
    ovol* v = c->vs;
    if ((uint32_t)v->incident != VOL_AIR)
    {
        col = cmul(col, beer(R, v));
        if (c->update)
        {
            v->distance = 0.0f;
            if (v->inside)
                vol_pop(v, 1);
        }
    }
    *pdf = fabsf(ll.z * M_INV_PI_F);
    return col;
}

/* The principled BSDF, layer by layer (Principled.h:955-1512, :1620-1800). */
static rgb principled_eval(const rctx* R, const omat* m, const bctx* c, float* pdf)
{
    *pdf = 0.0f;
    int outside = !c->vs->inside;
    int refr = vdot(c->sn, c->L) < 0.0f && outside;
    frames f = make_frames(m, c);
    const float inc = medium_ior(R, c->vs->incident);
    lobes w = weights(m, outside), p = probas(w);
    rgb layers = C1(1.0f), out = C1(0.0f);
    /* coat */
    if (w.coat > 0.0f && ((f.lv.z > 0.0f && f.ll.z > 0.0f) || refr))
    {
        float cp = 0.0f;
        rgb contrib = C1(0.0f);
        if (!refr)
        {
            float HoL = hclamp(1.0e-8f, 1.0f, vdot(f.lh, f.ll));
            rgb F = C1(fresnel_d2(HoL, inc, m->coat_ior));
            contrib = ggx_dielectric(R, m, c, m->coat_roughness, m->coat_anisotropy, F, f.lv, f.ll, f.lh, &cp);
            contrib = cmul(cscale(contrib, w.coat), layers);
        }
        *pdf += cp * p.coat;
        rgb below = C1(1.0f);
        below = cscale(below, 1.0f - fresnel_d2(fabsf(f.ll.z), inc, m->coat_ior));
        float vf = fresnel_d2(fabsf(f.lv.z), inc, m->coat_ior);
        below = cscale(below, 1.0f - vf);
        if (!(m->coat_absorption.r == 1.0f && m->coat_absorption.g == 1.0f && m->coat_absorption.b == 1.0f))
        {
            float ci = m->coat_ior;
            float ia = hmax(1.0e-6f, sqrtf(1.0f - (1.0f - f.ll.z * f.ll.z) / (ci * ci)));
            float oa = hmax(1.0e-6f, sqrtf(1.0f - (1.0f - f.lv.z * f.lv.z) / (ci * ci)));
            float dist = 1.0f / ia + 1.0f / oa;
            rgb a = m->coat_absorption;
            rgb pw = C3(orc_powf(sqrtf(a.r), dist), orc_powf(sqrtf(a.g), dist), orc_powf(sqrtf(a.b), dist));
            rgb e = cscale(csub(C1(0.0f), csub(C1(1.0f), pw)), m->coat_thickness);
            below = cmul(below, C3(orc_expf(e.r), orc_expf(e.g), orc_expf(e.b)));
        }
        below = cmul(below, coat_darkening(m, m->coat_ior / inc, vf));
        below = clerp(C1(1.0f), below, m->coat);
        layers = cmul(layers, below);
        out = cadd(out, contrib);
    }
    else
        out = cadd(out, C1(0.0f));
    /* sheen */
    if ((w.sheen > 0.0f && f.lv.z > 0.0f && f.ll.z > 0.0f) || refr)
    {
        float sp, refl;
        rgb contrib = sheen_eval(R, m, f.lv, f.ll, &sp, &refl);
        contrib = cmul(cscale(contrib, w.sheen), layers);
        *pdf += sp * p.sheen;
        layers = cscale(layers, 1.0f - m->sheen * refl);
        out = cadd(out, contrib);
    }
    else
        out = cadd(out, C1(0.0f));
    /* metal lobes */
    float wm[2] = {w.m1 * (float)!refr, w.m2 * (float)!refr};
    float pm[2] = {p.m1, p.m2}, rm[2] = {m->roughness, m->second_roughness};
    for (int k = 0; k < 2; k++)
    {
        rgb contrib = C1(0.0f);
        if (wm[k] > 0.0f && f.lvr.z > 0.0f && f.llr.z > 0.0f)
        {
            float mp = 0.0f;
            contrib = metallic_eval(R, m, c, rm[k], inc, f.lvr, f.llr, f.lhr, &mp);
            contrib = cmul(cscale(contrib, wm[k]), layers);
            *pdf += mp * pm[k];
        }
        out = cadd(out, contrib);
    }
    /* glass (rotated frame) */
    if (w.glass > 0.0f)
    {
        float gp = 0.0f;
        rgb contrib = cmul(cscale(glass_eval(R, m, c, f.lvr, f.llr, &gp), w.glass), layers);
        *pdf += gp * p.glass;
        out = cadd(out, contrib);
    }
    /* glossy base: specular over diffuse */
    rgb glossy = C1(0.0f);
    if (w.spec > 0.0f && ((f.lvr.z > 0.0f && f.llr.z > 0.0f) || (refr && c->info == FROM_DIFFUSE_TRANSMISSION)))
    {
        float rel = spec_rel_ior(m, inc);
        float sp = 0.0f;
        rgb F = spec_fresnel(m, rel, vdot(f.llr, f.lhr));
        rgb contrib = ggx_dielectric(R, m, c, m->roughness, m->anisotropy, F, f.lvr, f.llr, f.lhr, &sp);
        contrib = cmul(contrib, clerp(C1(1.0f), cscale(m->specular_color, m->specular_tint), m->specular));
        contrib = cmul(cscale(contrib, w.spec), layers);
        rgb below = C1(1.0f);
        rgb lf = cscale(spec_fresnel(m, rel, fabsf(f.llr.z)), 1.0f - m->diffuse_transmission);
        below = cmul(below, csub(C1(1.0f), lf));
        rgb vfc = spec_fresnel(m, rel, fabsf(f.lvr.z));
        below = cmul(below, csub(C1(1.0f), vfc));
        below = cmul(below, spec_darkening(m, rel, lum(vfc)));
        below = clerp(C1(1.0f), below, m->specular);
        layers = cmul(layers, below);
        *pdf += sp * p.spec;
        glossy = cadd(glossy, contrib);
    }
    else
        glossy = cadd(glossy, C1(0.0f));
    float wd = w.diff * (float)!refr;
    if (wd > 0.0f && f.lv.z > 0.0f && f.ll.z > 0.0f)
    {
        float dp = f.ll.z * M_INV_PI_F;
        rgb contrib = cmul(cscale(cscale(m->base_color, M_INV_PI_F), wd), layers);
        *pdf += dp * p.diff;
        glossy = cadd(glossy, contrib);
    }
    out = cadd(out, cdivs(glossy, ec_glossy(R, m, inc, f.lv.z)));
    /* diffuse transmission */
    if (w.dt > 0.0f && f.ll.z < 0.0f)
    {
        float dp = 0.0f;
        rgb contrib = cmul(cscale(dt_eval(R, m, c, f.lv, f.ll, &dp), w.dt), layers);
        *pdf += dp * p.dt;
        out = cadd(out, contrib);
    }
    return cdivs(out, ec_clearcoat(R, m, inc, f.lv.z));
}

static float principled_pdf(const rctx* R, const omat* m, const bctx* c)
{
    float pdf = 0.0f;
    int outside = !c->vs->inside;
    int refr = vdot(c->sn, c->L) < 0.0f && outside;
    frames f = make_frames(m, c);
    lobes w = weights(m, outside), p = probas(w);
    if (w.coat > 0.0f && ((f.lv.z > 0.0f && f.ll.z > 0.0f) || refr))
        pdf += (refr ? 0.0f : ggx_dielectric_pdf(R, m, c, m->coat_roughness, m->coat_anisotropy, f.lv, f.ll, f.lh)) *
               p.coat;
    if ((w.sheen > 0.0f && f.lv.z > 0.0f && f.ll.z > 0.0f) || refr)
        pdf += sheen_pdf(R, m, f.lv, f.ll) * p.sheen;
    if (w.m1 * (float)!refr > 0.0f && f.lvr.z > 0.0f && f.llr.z > 0.0f)
        pdf += metal_pdf(R, m, c, m->roughness, f.lvr, f.llr, f.lhr) * p.m1;
    if (w.m2 * (float)!refr > 0.0f && f.lvr.z > 0.0f && f.llr.z > 0.0f)
        pdf += metal_pdf(R, m, c, m->second_roughness, f.lvr, f.llr, f.lhr) * p.m2;
    if (w.glass > 0.0f)
        pdf += glass_pdf(R, m, c, f.lvr, f.llr) * p.glass;
    float glossy = 0.0f;
    if (w.spec > 0.0f && ((f.lvr.z > 0.0f && f.llr.z > 0.0f) || (refr && c->info == FROM_DIFFUSE_TRANSMISSION)))
        glossy += ggx_dielectric_pdf(R, m, c, m->roughness, m->anisotropy, f.lvr, f.llr, f.lhr) * p.spec;
    if (w.diff * (float)!refr > 0.0f && f.lv.z > 0.0f && f.ll.z > 0.0f)
        glossy += (f.ll.z > 0.0f ? f.ll.z * M_INV_PI_F : 0.0f) * p.diff;
    pdf += glossy;
    if (w.dt > 0.0f && f.ll.z < 0.0f)
        pdf += (f.lv.z * f.ll.z > 0.0f ? 0.0f : fabsf(f.ll.z * M_INV_PI_F)) * p.dt;
    return pdf;
}

static rgb principled_sample(const rctx* R, const omat* m, bctx* c, v3* dir, float* pdf, uint32_t* st)
{
    *pdf = 0.0f;
    lobes w = weights(m, !c->vs->inside), p = probas(w);
    float cdf0 = p.coat, cdf1 = cdf0 + p.sheen, cdf2 = cdf1 + p.m1, cdf3 = cdf2 + p.m2;
    float cdf4 = cdf3 + p.spec, cdf5 = cdf4 + p.diff, cdf6 = cdf5 + p.dt;
    float r1 = orc_rand01(st);
    int s_dt = r1 > cdf5 && r1 < cdf6, s_glass = r1 > cdf6;
    if (c->update && !s_glass && !s_dt)
        vol_pop(c->vs, 0);
    v3 TR, BR;
    build_rotated_onb(c->sn, &TR, &BR, m->anisotropy_rotation * M_PI_F);
    v3 lvr = to_local(TR, BR, c->sn, c->view);
    if (r1 < cdf0)
    {
        v3 TC, BC;
        build_rotated_onb(c->sn, &TC, &BC, m->coat_anisotropy_rotation * M_PI_F);
        v3 lvc = to_local(TC, BC, c->sn, c->view);
        c->info = FROM_COAT;
        *dir = to_world(TC, BC, c->sn, ggx_sample(regularize(R, c->reg, m->coat_roughness, c->path_r), m->coat_anisotropy, lvc, st));
    }
    else if (r1 < cdf1)
    {
        v3 T, B;
        build_onb(c->sn, &T, &B);
        v3 lv = to_local(T, B, c->sn, c->view);
        *dir = to_world(T, B, c->sn, sheen_sample(R, m, lv, st));
    }
    else if (r1 < cdf2)
    {
        c->info = FROM_FIRST_METAL;
        *dir = to_world(TR, BR, c->sn, ggx_sample(regularize(R, c->reg, m->roughness, c->path_r), m->anisotropy, lvr, st));
    }
    else if (r1 < cdf3)
    {
        c->info = FROM_SECOND_METAL;
        *dir = to_world(TR, BR, c->sn, ggx_sample(regularize(R, c->reg, m->second_roughness, c->path_r), m->anisotropy, lvr, st));
    }
    else if (r1 < cdf4)
    {
        c->info = FROM_SPECULAR;
        *dir = to_world(TR, BR, c->sn, ggx_sample(regularize(R, c->reg, m->roughness, c->path_r), m->anisotropy, lvr, st));
    }
    else if (r1 < cdf5)
    {
        c->info = FROM_DIFFUSE;
        *dir = cosine_around_normal(c->sn, st);
    }
    else if (r1 < cdf6)
    {
        c->info = FROM_DIFFUSE_TRANSMISSION;
        *dir = cosine_around_normal(vneg(c->sn), st);
    }
    else
        *dir = to_world(TR, BR, c->sn, glass_sample(R, m, c, lvr, st));
    if (vdot(*dir, c->gn) < 0.0f && !s_glass && !s_dt)
        return C1(0.0f);
    c->L = *dir;
    return principled_eval(R, m, c, pdf);
}

/* ------------------------------------------------------------------ */
/* Octahedral normals (Packing.h:237-360)                               */
/* ------------------------------------------------------------------ */
static float snz(float k) { return k >= 0.0f ? 1.0f : -1.0f; }

static uint32_t oct_pack(v3 v)
{
    float l1 = 1.0f / (fabsf(v.x) + fabsf(v.y) + fabsf(v.z));
    float ex = v.x * l1, ey = v.y * l1;
    if (v.z < 0.0f)
    {
        float nx = (1.0f - fabsf(ey)) * snz(ex), ny = (1.0f - fabsf(ex)) * snz(ey);
        ex = nx;
        ey = ny;
    }
    float sx = roundf(hclamp(0.0f, 2.0f, ex + 1.0f) * 2047.0f);
    float sy = roundf(hclamp(0.0f, 2.0f, ey + 1.0f) * 2047.0f);
    float ux = sx / 16.0f, t = floorf(sy / 256.0f);
    float uy = ((ux - floorf(ux)) * 256.0f) + t, uz = sy - (t * 256.0f);
    return (uint32_t)(unsigned char)ux | ((uint32_t)(unsigned char)uy << 8) | ((uint32_t)(unsigned char)uz << 16);
}

static v3 oct_unpack(uint32_t p)
{
    unsigned bx = p & 0xff, by = (p >> 8) & 0xff, bz = (p >> 16) & 0xff;
    float temp = (float)by / 16.0f;
    float sx = (float)bx * 16.0f + floorf(temp);
    float sy = (temp - floorf(temp)) * 256.0f * 16.0f + (float)bz;
    float x = hclamp(-1.0f, 1.0f, (sx / 2047.0f) - 1.0f), y = hclamp(-1.0f, 1.0f, (sy / 2047.0f) - 1.0f);
    v3 v = V(x, y, 1.0f - fabsf(x) - fabsf(y));
    if (v.z < 0.0f)
    {
        float tx = v.x, ty = v.y;
        v.x = (1.0f - fabsf(ty)) * snz(tx);
        v.y = (1.0f - fabsf(tx)) * snz(ty);
    }
    return vnorm(v);
}

uint32_t orc_oct_pack(const float* n) { return oct_pack(V(n[0], n[1], n[2])); }
void orc_oct_unpack(uint32_t p, float* out) { v3 v = oct_unpack(p); out[0] = v.x; out[1] = v.y; out[2] = v.z; }

/* ------------------------------------------------------------------ */
/* Hit attributes + back-face fix (Intersect.h:142-176, 222-245)        */
/* ------------------------------------------------------------------ */
typedef struct { v3 p, gn, sn; int32_t prim; } hitinfo;

static void fix_backfacing(v3 view, v3* gn, v3* sn)
{
    if (vdot(view, *gn) < 0.0f)
    {
        *gn = vscale(*gn, -1.0f);
        *sn = vscale(*sn, -1.0f);
    }
    if (vdot(view, *sn) < 0.0f)
        *sn = vscale(*sn, -1.0f);
    v3 refl = reflect(view, *sn);
    if (vdot(refl, *gn) <= 0.0f)
    {
        refl = vsub(refl, vnorm(vscale(*gn, vdot(refl, *gn) - 0.01f)));
        *sn = vnorm(vadd(view, refl));
    }
}

/* get_shading_normal (Intersect.h:114-133): with a vertex normal at the
 * triangle's first vertex, normalize(n_B u + n_C v + n_A (1 - u - v))
 * (uv_interpolate, Texture.h:248-251); otherwise the geometric normal. */
static v3 shading_normal(const rctx* R, int32_t prim, float u, float v, v3 gn)
{
    if (!R->has_normals)
        return gn;
    const int32_t* t = R->idx + 3 * (int64_t)prim;
    if (!R->has_normals[t[0]])
        return gn;
    const float* na = R->vnormals + 3 * (int64_t)t[0];
    const float* nb = R->vnormals + 3 * (int64_t)t[1];
    const float* nc = R->vnormals + 3 * (int64_t)t[2];
    float w = (1.0f - u) - v;
    v3 n = vadd(vadd(vscale(V(nb[0], nb[1], nb[2]), u), vscale(V(nc[0], nc[1], nc[2]), v)),
                vscale(V(na[0], na[1], na[2]), w));
    return vnorm(n);
}

void orc_shading_normal(const int32_t* idx, const float* vertex_normals, const uint8_t* has_vertex_normals,
                        int32_t prim, float u, float v, const float* geometric_normal, float* out)
{
    rctx R;
    memset(&R, 0, sizeof(R));
    R.idx = idx;
    R.has_normals = has_vertex_normals;
    R.vnormals = vertex_normals;
    v3 n = shading_normal(&R, prim, u, v, V(geometric_normal[0], geometric_normal[1], geometric_normal[2]));
    out[0] = n.x;
    out[1] = n.y;
    out[2] = n.z;
}

static int trace_main(const rctx* R, v3 o, v3 d, int32_t last, hitinfo* out)
{
    orc_hit h = trace1(R, o, d, last, INFINITY);
    if (h.prim < 0)
        return 0;
    v3 a, b, c;
    tri(R, h.prim, &a, &b, &c);
    out->p = vadd(o, vscale(d, h.t));
    out->gn = vnorm(vnorm(vcross(vsub(b, a), vsub(c, a))));
    out->sn = shading_normal(R, h.prim, h.u, h.v, out->gn);
    out->prim = h.prim;
    fix_backfacing(vneg(d), &out->gn, &out->sn);
    return 1;
}

/* trace_main_path_ray (Intersect.h:200-275): every boundary hit is pushed
 * on the interior stack; one of lower priority than the medium the ray is
 * in is skipped and the ray goes on from the hit point (same filter); the
 * distance inside a medium is accumulated (twice over a skipped boundary,
 * as there). Reaching a dispersive glass samples the wavelength. */
static int trace_path(const rctx* R, v3 o, v3 d, int32_t last, hitinfo* out, ovol* v, uint32_t* st)
{
    int skip;
    do
    {
        orc_hit h = trace1(R, o, d, last, INFINITY);
        if (h.prim < 0)
            return 0;
        v3 a, b, c;
        tri(R, h.prim, &a, &b, &c);
        out->p = vadd(o, vscale(d, h.t));
        out->gn = vnorm(vnorm(vcross(vsub(b, a), vsub(c, a))));
        out->sn = shading_normal(R, h.prim, h.u, h.v, out->gn);
        out->prim = h.prim;
        int mat = R->mat_idx[h.prim];
        skip = vol_push(v, mat, R->mats[mat].dielectric_priority);
        if (v->inside)
            v->distance += h.t;
        fix_backfacing(vneg(d), &out->gn, &out->sn);
        if (skip)
        {
            o = out->p;
            v->distance += h.t;
        }
    } while (skip);
    const omat* m = &R->mats[R->mat_idx[out->prim]];
    if (m->dispersion_scale > 0.0f && m->specular_transmission > 0.0f && v->wavelength == 0.0f)
        v->wavelength = -(orc_rand01(st) * (830.0f - 360.0f) + 360.0f);
    return 1;
}

/* wavelength_to_RGB_fit (Dispersion.h:433-458) */
static rgb wavelength_rgb(float w)
{
    rgb c;
    if (w < 463.0f)
        c.r = -1.2776028240727566e-01f / (1.0f + orc_expf((w - 4.2680623367293401e+02f) / 8.2197460736637176e+00f)) +
              -1.3925673552505122e-11f * orc_expf((w - 45.0f) / 1.8175459086411596e+01f) + 1.2898689750552100e-01f;
    else if (w > 553.0f)
    {
        float t = (w - 6.0625724092824566e+02f) * (1.0f / 2.6577826611702449e+01f);
        c.r = 1.7963649137825513e+01f * (1.0f / 2.6577826611702449e+01f) * orc_expf(-0.5f * (t * t)) +
              2.5574660155104657e-03f;
    }
    else
        c.r = 0.0f;
    float tg = (w - 5.4209217455705152e+02f) / -2.9598170255834638e+01f;
    c.g = 3.4962267376163049e+02f * orc_expf(-0.5f * (tg * tg)) / w;
    c.b = orc_expf(3.2987659944421112e+03f + (-2.0975839709372405e+05f / w) - 4.6368268395094020e+02f * orc_logf(w));
    return cscale(c, 10.0f);
}

/* get_dispersion_ray_color (Dispersion.h:512-535) */
static rgb dispersion_filter(float* w, float scale)
{
    if (scale == 0.0f || *w >= 0.0f)
        return C1(1.0f);
    *w = -*w;
    return wavelength_rgb(*w);
}

/* ------------------------------------------------------------------ */
/* Emissive-triangle sampling (LightUtils.h:51-76, :142-309,            */
/* AliasTable.h:13-22) and NEE with MIS (Lights.h:123-226,              */
/* PDFConversion.h:19-38, Sampling.h:103-136)                           */
/* ------------------------------------------------------------------ */

typedef struct { int32_t tri; v3 normal; float area; rgb emission; v3 point; float area_pdf; } olight;

static olight sample_emissive_with(const rctx* R, int32_t strategy, uint32_t* st)
{
    olight L;
    int32_t e = orc_random_index(st, R->ne);
    if (strategy == 1)
    {
        float p = R->alias_probas[e];
        if (orc_rand01(st) > p)
            e = R->alias_alias[e];
    }
    L.tri = R->emissive[e];
    v3 A, B, Cc;
    tri(R, L.tri, &A, &B, &Cc);
    float r1 = orc_rand01(st), r2 = orc_rand01(st);
    if (r2 > r1) { r1 *= 0.5f; r2 -= r1; }
    else { r2 *= 0.5f; r1 -= r2; }
    v3 AB = vsub(B, A), AC = vsub(Cc, A);
    v3 n = vcross(AB, AC);
    float len = vlen(n);
    L.point = vadd(vadd(A, vscale(AB, r1)), vscale(AC, r2));
    L.normal = vdivs(n, len);
    L.area = 0.5f * len;
    L.emission = R->mats[R->mat_idx[L.tri]].emission;
    L.area_pdf = 1.0f / L.area;
    if (strategy == 1)
        L.area_pdf *= (lum(L.emission) * L.area) / R->alias_sum;
    else
        L.area_pdf /= (float)R->ne;
    return L;
}

/* ------------------------------------------------------------------ */
/* Ambient light: R9G9B9E5 texels (Packing.h:441-485) and the envmap   */
/* (Envmap.h:29-276, Texture.h:209-240, PathTracing.h:121-170)          */
/* ------------------------------------------------------------------ */

/* RGBE9995Packed::pack: the channels are clamped to [0, 511/256 2^15]; a
 * power of two whose exponent is the largest channel's (floored at 2^-16)
 * plus 15 -- plus 16 when that channel's top nine significant bits round
 * up -- is added to each channel, so the float addition rounds each to nine
 * bits below that power; those low nine bits of the sums and the power's
 * exponent + 2 (mod 32) make the 32-bit word. */
uint32_t orc_rgb9e5_pack(float r, float g, float b)
{
    const float hi = 65408.0f, lo = 1.52587890625e-05f; /* 511/256 2^15, 2^-16 */
    r = hmin(hi, hmax(0.0f, r));
    g = hmin(hi, hmax(0.0f, g));
    b = hmin(hi, hmax(0.0f, b));
    float m = hmax(hmax(lo, r), hmax(g, b));
    uint32_t mb;
    memcpy(&mb, &m, 4);
    uint32_t bias_bits = (mb + (15u << 23) + (1u << 14)) & 0x7f800000u;
    float bias;
    memcpy(&bias, &bias_bits, 4);
    float sr = r + bias, sg = g + bias, sb = b + bias;
    uint32_t R9, G9, B9;
    memcpy(&R9, &sr, 4);
    memcpy(&G9, &sg, 4);
    memcpy(&B9, &sb, 4);
    uint32_t e = ((bias_bits >> 23) + 2u) & 31u;
    return (e << 27) | ((B9 & 0x1ffu) << 18) | ((G9 & 0x1ffu) << 9) | (R9 & 0x1ffu);
}

void orc_rgb9e5_unpack(uint32_t p, float* rgb)
{
    int ex = (int)(p >> 27) - 24;
    rgb[0] = ldexpf((float)(p & 0x1ffu), ex);
    rgb[1] = ldexpf((float)((p >> 9) & 0x1ffu), ex);
    rgb[2] = ldexpf((float)((p >> 18) & 0x1ffu), ex);
}

static int env_on(const rctx* R) { return R->env_type == 2 && R->env_intensity > 0.0f && R->env_w > 0; }

/* matrix_X_vec (Math.h:636-649) */
static v3 env_rot(const float m[3][3], v3 v)
{
    return V(m[0][0] * v.x + m[1][0] * v.y + m[2][0] * v.z, m[0][1] * v.x + m[1][1] * v.y + m[2][1] * v.z,
             m[0][2] * v.x + m[1][2] * v.y + m[2][2] * v.z);
}

static int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* sample_environment_map_texture, nearest texel (Texture.h:234-239); the
 * texel coordinates are kept in the image */
static rgb env_texel(const rctx* R, float u, float v)
{
    float fx = u * (float)(uint32_t)(R->env_w - 1), fy = v * (float)(uint32_t)(R->env_h - 1);
    int32_t x = isnan(fx) ? 0 : clampi((int32_t)hmax(-1.0f, hmin(fx, (float)R->env_w)), 0, R->env_w - 1);
    int32_t y = isnan(fy) ? 0 : clampi((int32_t)hmax(-1.0f, hmin(fy, (float)R->env_h)), 0, R->env_h - 1);
    float c[3];
    orc_rgb9e5_unpack(R->env_texels[x + y * R->env_w], c);
    return cscale(C3(c[0], c[1], c[2]), R->env_intensity);
}

/* eval_envmap_no_pdf (Envmap.h:29-38) */
static rgb env_radiance(const rctx* R, v3 dir)
{
    v3 d = env_rot(R->env_to_env, dir);
    float u = 0.5f + orc_atan2f(d.z, d.x) * 0.15915494309189533577f;
    float v = 0.5f + orc_asinf(d.y) * 0.31830988618379067154f;
    return env_texel(R, u, v);
}

/* Envmap.h:122-131 / :155-163: luminance pdf over the texels, in solid angle */
static float env_pdf(const rctx* R, rgb radiance, float sin_theta)
{
    float pdf = lum(radiance) / (R->env_sum * R->env_intensity);
    pdf *= (float)((uint32_t)R->env_w * (uint32_t)R->env_h);
    return pdf / (19.73920880217871723767f * sin_theta);
}

/* envmap_sample (Envmap.h:77-133) with DeviceAliasTable::sample */
static rgb env_sample(const rctx* R, v3* dir, float* pdf, uint32_t* st)
{
    int32_t i = orc_random_index(st, R->env_w * R->env_h);
    if (orc_rand01(st) > R->env_probas[i])
        i = R->env_alias[i];
    int32_t y = (int32_t)floorf((float)i / (float)(uint32_t)R->env_w);
    int32_t x = (int32_t)floorf((float)i - (float)y * (float)(uint32_t)R->env_w);
    float u = (float)x / (float)(uint32_t)R->env_w;
    float v = (float)y / (float)(uint32_t)R->env_h;
    float phi = u * 6.28318530717958647693f;
    float theta = hmax(1.0e-5f, v * 3.14159265358979323846f);
    float ct = orc_cosf(theta), s_t = orc_sinf(theta);
    *dir = env_rot(R->env_to_world, V(-s_t * orc_cosf(phi), -ct, -s_t * orc_sinf(phi)));
    rgb rad = env_texel(R, u, v);
    *pdf = env_pdf(R, rad, s_t);
    return rad;
}

/* envmap_eval (Envmap.h:139-166) */
static rgb env_eval(const rctx* R, v3 dir, float* pdf)
{
    rgb rad = env_radiance(R, dir);
    float theta = orc_acosf(-dir.y);
    *pdf = env_pdf(R, rad, orc_sinf(theta));
    return rad;
}

/* path_tracing_miss_gather_envmap (PathTracing.h:121-170), clamps off */
static rgb env_miss(const rctx* R, v3 dir, int bounce)
{
    if (R->env_type == 1)
        return C3(R->env_uniform[0], R->env_uniform[1], R->env_uniform[2]);
    if (R->env_type != 2 || R->env_intensity == 0.0f || R->env_w == 0 || bounce != 0)
        return C1(0.0f);
    rgb c = env_radiance(R, dir);
    if (!R->env_scale_bg)
        c = cdivs(c, R->env_intensity);
    return c;
}

static float bal2(float a, float b) { return a == 0.0f ? 0.0f : a / (1.0f * a + 1.0f * b); }

/* sample_environment_map_with_mis (Envmap.h:168-255): the envmap light
 * sample, then a BSDF sample for MIS, each with its shadow ray */
static rgb sample_envmap_mis(const rctx* R, const omat* m, const hitinfo* hi, v3 view, float path_r, uint32_t* st,
                             ovol* vs)
{
    v3 edir;
    float epdf;
    rgb ecol = env_sample(R, &edir, &epdf, st);
    rgb env_mis = C1(0.0f);
    float ecos = vdot(hi->sn, edir);
    if (epdf > 0.0f && ecos > 0.0f)
    {
        orc_hit sh = trace1(R, hi->p, edir, hi->prim, INFINITY);
        if (!(sh.prim >= 0 && sh.t < 1.0e35f - 1.0e-4f))
        {
            bctx c = {view, hi->sn, hi->gn, edir, NO_INFO, REG_MIS, path_r, vs, 0};
            float bpdf;
            rgb f = principled_eval(R, m, &c, &bpdf);
            float w = bal2(epdf, bpdf);
            env_mis = cdivs(cdivs(cmul(cscale(cscale(f, ecos), w), ecol), epdf), 1.0f);
        }
    }
    rgb bsdf_mis = C1(0.0f);
    bctx c = {view, hi->sn, hi->gn, V(0, 0, 0), NO_INFO, REG_MIS, path_r, vs, 0};
    v3 dir;
    float bpdf;
    rgb f = principled_sample(R, m, &c, &dir, &bpdf, st);
    float bcos = fabsf(vdot(hi->sn, dir));
    if (bpdf > 0.0f)
    {
        orc_hit sh = trace1(R, hi->p, dir, hi->prim, INFINITY);
        if (!(sh.prim >= 0 && sh.t < 1.0e35f - 1.0e-4f))
        {
            float ep;
            rgb rad = env_eval(R, dir, &ep);
            if (ep > 0.0f)
            {
                float w = bal2(bpdf, ep);
                bsdf_mis = cdivs(cmul(cscale(cscale(rad, w), bcos), f), bpdf);
            }
        }
    }
    return cadd(bsdf_mis, env_mis);
}

static olight sample_emissive(const rctx* R, uint32_t* st)
{
    return sample_emissive_with(R, R->s.direct_light_sampling_base_strategy, st);
}

static float emissive_hit_area_pdf_with(const rctx* R, int32_t strategy, float area, rgb emission)
{
    float pdf = 1.0f / area;
    if (strategy == 1)
        pdf *= (lum(emission) * area) / R->alias_sum;
    else
        pdf /= (float)R->ne;
    return pdf;
}

static float emissive_hit_area_pdf(const rctx* R, float area, rgb emission)
{
    return emissive_hit_area_pdf_with(R, R->s.direct_light_sampling_base_strategy, area, emission);
}

static rgb sample_one_light_mis(const rctx* R, const omat* m, const hitinfo* hi, v3 view, float path_r, uint32_t* st,
                                ovol* vs)
{
    rgb light_mis = C1(0.0f);
    {
        olight L = sample_emissive(R, st);
        v3 point = L.point, ln = L.normal;
        float apdf = L.area_pdf;
        rgb emission = L.emission;
        if (apdf > 0.0f)
        {
            v3 to = vsub(point, hi->p);
            float dist = vlen(to);
            v3 dn = vdivs(to, dist);
            /* evaluate_shadow_ray_occluded, CPU branch (Intersect.h:297-323) */
            orc_hit sh = trace1(R, hi->p, dn, hi->prim, INFINITY);
            int in_shadow = sh.prim >= 0 && sh.t < dist - 1.0e-4f;
            if (!in_shadow)
            {
                bctx c = {view, hi->sn, hi->gn, dn, NO_INFO, REG_MIS, path_r, vs, 0};
                float bpdf;
                rgb f = principled_eval(R, m, &c, &bpdf);
                if (bpdf > 0.0f)
                {
                    float cl = hmax(0.0f, vdot(ln, vneg(dn)));
                    if (cl > 1.0e-5f)
                    {
                        float lpdf = cl < 1.0e-8f ? 0.0f : apdf * (dist * dist) / cl;
                        float w = lpdf == 0.0f ? 0.0f : lpdf / (1.0f * lpdf + 1.0f * bpdf);
                        float cosine = fabsf(vdot(hi->sn, dn));
                        light_mis = cdivs(cdivs(cscale(cmul(cscale(f, cosine), emission), w), lpdf), 1.0f);
                    }
                }
            }
        }
    }
    rgb bsdf_mis = C1(0.0f);
    {
        bctx c = {view, hi->sn, hi->gn, V(0, 0, 0), NO_INFO, REG_MIS, path_r, vs, 0};
        v3 dir;
        float bpdf;
        rgb f = principled_sample(R, m, &c, &dir, &bpdf, st);
        if (bpdf > 0.0f)
        {
            /* evaluate_bsdf_light_sample_ray, CPU branch (Intersect.h:484-535) */
            orc_hit h = trace1(R, hi->p, dir, hi->prim, INFINITY);
            if (h.prim >= 0 && h.t < 1.0e35f - 1.0e-4f)
            {
                rgb em = R->mats[R->mat_idx[h.prim]].emission;
                v3 a, b, cc;
                tri(R, h.prim, &a, &b, &cc);
                v3 hgn = vnorm(vnorm(vcross(vsub(b, a), vsub(cc, a))));
                float cl = hmax(0.0f, vdot(hgn, vneg(dir)));
                if (!cblack(em) && cl > 0.0f)
                {
                    float pa = emissive_hit_area_pdf(R, R->areas[h.prim], em);
                    float lpdf = cl < 1.0e-8f ? 0.0f : pa * (h.t * h.t) / cl;
                    float w = bpdf == 0.0f ? 0.0f : bpdf / (1.0f * bpdf + 1.0f * lpdf);
                    float cosine = fabsf(vdot(hi->sn, dir));
                    bsdf_mis = cdivs(cscale(cmul(cscale(f, cosine), em), w), bpdf);
                }
            }
        }
    }
    return cadd(light_mis, bsdf_mis);
}

/* Sampling.h:103-116 */
static float balance_heuristic(float pa, float na, float pb, float nb)
{
    return pa == 0.0f ? 0.0f : pa / (na * pa + nb * pb);
}

/* sample_lights_RIS (RIS/RIS.h:83-258): light candidates, then the BSDF
 * candidate, streamed through one reservoir (RIS_Reservoir.h:29-45: every
 * candidate draws its selection number, selected when u < w / wsum), then the
 * kept sample shaded (RIS.h:18-80). can_do_light_sampling is true with
 * microfacet regularisation (MaterialUtils.h:71-78). The candidates' BSDF
 * evaluations take the path's volume state by reference, as in the reference. */
static rgb sample_lights_ris(const rctx* R, const omat* m, const hitinfo* hi, v3 view, float path_r, uint32_t* st,
                             ovol* vs)
{
    const int32_t nlc = R->s.ris_number_of_light_candidates, nbc = R->s.ris_number_of_bsdf_candidates;
    const float nl = (float)nlc, nb = (float)nbc;
    float wsum = 0.0f;
    /* the kept sample */
    int is_bsdf = 0;
    v3 point = V(0, 0, 0);
    rgb emission = C1(0.0f), fb = C1(0.0f);
    float target = 0.0f, cos_b = 0.0f;
    for (int32_t i = 0; i < nlc; i++)
    {
        olight L = sample_emissive(R, st);
        float tf = 0.0f, w = 0.0f;
        if (L.area_pdf > 0.0f)
        {
            v3 to = vsub(L.point, hi->p);
            float dist = vlen(to);
            v3 dn = vdivs(to, dist);
            float cl = hmax(0.0f, vdot(L.normal, vneg(dn)));
            float cp = fabsf(vdot(hi->sn, dn));
            if (cp > 0.0f && cl > 1.0e-6f)
            {
                bctx c = {view, hi->sn, hi->gn, dn, NO_INFO, REG_MIS, path_r, vs, 0};
                float bpdf = 0.0f;
                rgb f = principled_eval(R, m, &c, &bpdf);
                tf = lum(cscale(cmul(f, L.emission), cp));
                float spdf = cl < 1.0e-8f ? 0.0f : L.area_pdf * (dist * dist) / cl;
                w = balance_heuristic(spdf, nl, bpdf, nb) * tf / spdf;
            }
        }
        wsum += w;
        if (orc_rand01(st) < w / wsum)
        {
            is_bsdf = 0;
            point = L.point;
            emission = L.emission;
            target = tf;
        }
    }
    for (int32_t i = 0; i < nbc; i++)
    {
        bctx c = {view, hi->sn, hi->gn, V(0, 0, 0), NO_INFO, REG_MIS, path_r, vs, 0};
        v3 dir;
        float bpdf = 0.0f, w = 0.0f, tf = 0.0f, cp = 0.0f;
        rgb f = principled_sample(R, m, &c, &dir, &bpdf, st);
        rgb em = C1(0.0f);
        int hit_light = 0;
        if (bpdf > 0.0f)
        {
            /* evaluate_bsdf_light_sample_ray, CPU branch (Intersect.h:484-535) */
            orc_hit h = trace1(R, hi->p, dir, hi->prim, INFINITY);
            if (h.prim >= 0 && h.t < 1.0e35f - 1.0e-4f)
            {
                em = R->mats[R->mat_idx[h.prim]].emission;
                v3 a, b, cc;
                tri(R, h.prim, &a, &b, &cc);
                v3 hgn = vnorm(vnorm(vcross(vsub(b, a), vsub(cc, a))));
                float cl = hmax(0.0f, vdot(hgn, vneg(dir)));
                if (!cblack(em) && cl > 0.0f)
                {
                    cp = fabsf(vdot(hi->sn, dir));
                    tf = lum(cscale(cmul(f, em), cp));
                    float pa = emissive_hit_area_pdf(R, R->areas[h.prim], em);
                    float lpdf = cl < 1.0e-8f ? 0.0f : pa * (h.t * h.t) / cl;
                    w = balance_heuristic(bpdf, nb, lpdf, nl) * tf / bpdf;
                    hit_light = 1;
                }
            }
        }
        wsum += w;
        if (orc_rand01(st) < w / wsum)
        {
            is_bsdf = hit_light;
            emission = hit_light ? em : C1(0.0f);
            fb = f;
            cos_b = cp;
            target = tf;
        }
    }
    float ucw = wsum == 0.0f ? 0.0f : 1.0f / target * wsum;
    if (!(ucw > 0.0f))
        return C1(0.0f);
    if (is_bsdf)
        return cscale(cmul(cscale(fb, ucw), emission), cos_b);
    v3 to = vsub(point, hi->p);
    float dist = vlen(to);
    v3 dn = vdivs(to, dist);
    orc_hit sh = trace1(R, hi->p, dn, hi->prim, INFINITY);
    if (sh.prim >= 0 && sh.t < dist - 1.0e-4f)
        return C1(0.0f);
    bctx c = {view, hi->sn, hi->gn, dn, NO_INFO, REG_MIS, path_r, vs, 0};
    float bpdf;
    rgb f = principled_eval(R, m, &c, &bpdf);
    return cscale(cmul(cscale(f, ucw), emission), fabsf(vdot(hi->sn, dn)));
}

/* ------------------------------------------------------------------ */
/* ReSTIR DI reservoir (DI/Reservoir.h:52-187) and its final shading    */
/* (DI/FinalShading.h:16-125)                                           */
/* ------------------------------------------------------------------ */
typedef struct { v3 point; float target; int32_t M; int32_t tri; float ucw; uint32_t flags; float wsum; } ores;

#define RESTIR_UNOCCLUDED (1u << 7)
#define RESTIR_ENVMAP (1u << 0) /* RESTIR_DI_FLAGS_ENVMAP_SAMPLE (SampleFlags.h:16) */
#define RESTIR_LOBE_MASK (0x3fu << 1) /* flags_to_BSDF_incident_light_info keeps bits 1-6 */
#define FROM_GLASS_REFRACT_BIT (1u << 6)

/* ReSTIR_DI_get_light_sample_direction (DI/TargetFunction.h:13-29): an
 * envmap sample holds a direction in envmap space and lies at 1e35 */
static v3 sample_dir(const rctx* R, v3 p, v3 point, uint32_t flags, float* dist)
{
    if (flags & RESTIR_ENVMAP)
    {
        *dist = 1.0e35f;
        return env_rot(R->env_to_world, point);
    }
    v3 to = vsub(point, p);
    *dist = vlen(to);
    return vdivs(to, *dist);
}

static rgb restir_final_shading(const rctx* R, const omat* m, const hitinfo* hi, v3 view, size_t pix, ovol* vs)
{
    const ores* r = (const ores*)R->restir_out + pix;
    int env = (r->flags & RESTIR_ENVMAP) != 0;
    /* NaN weight sum with no selected sample; validate_reservoir
     * (FinalShading.h:103-108) for an envmap sample without the envmap */
    if (r->ucw <= 0.0f || (r->tri < 0 && !env) || (env && !env_on(R)))
        return C1(0.0f);
    float dist;
    v3 dir = sample_dir(R, hi->p, r->point, r->flags, &dist);
    int in_shadow = 0;
    if (r->flags & RESTIR_UNOCCLUDED)
        in_shadow = 0;
    else if (R->final_shading_visibility)
    {
        orc_hit h = trace1(R, hi->p, dir, hi->prim, INFINITY);
        in_shadow = h.prim >= 0 && h.t < dist - 1.0e-4f;
    }
    if (in_shadow)
        return C1(0.0f);
    bctx c = {view, hi->sn, hi->gn, dir, (int)(r->flags & RESTIR_LOBE_MASK), REG_MIS, 0.0f, vs, 0};
    float bpdf;
    rgb f = principled_eval(R, m, &c, &bpdf);
    float cosine = vdot(hi->sn, dir);
    if (r->flags & FROM_GLASS_REFRACT_BIT)
        cosine = fabsf(cosine);
    if (!(cosine > 0.0f))
        return C1(0.0f);
    rgb em;
    float conv = 1.0f;
    if (env)
    {
        float epdf;
        em = env_eval(R, dir, &epdf);
    }
    else
    {
        em = R->mats[R->mat_idx[r->tri]].emission;
        v3 a, b, cc;
        tri(R, r->tri, &a, &b, &cc);
        conv = hmax(0.0f, vdot(vnorm(vcross(vsub(b, a), vsub(cc, a))), vneg(dir)));
        conv /= dist * dist;
    }
    return cscale(cscale(cmul(cscale(f, r->ucw), em), cosine), conv);
}

/* ------------------------------------------------------------------ */
/* CameraRays + MegaKernel for one pixel                                */
/* ------------------------------------------------------------------ */
typedef struct { int32_t prim; v3 pos; uint32_t gn, sn; int32_t mat; } gbuf;

static void camera_rays_pixel(const rctx* R, int x, int y, uint32_t rn, gbuf* g)
{
    const int w = R->s.width, h = R->s.height;
    uint32_t pix = (uint32_t)x + (uint32_t)y * (uint32_t)w;
    uint32_t st = orc_pixel_seed(pix, R->sample_number, rn, R->s.freeze_random);
    float px = (float)x + 0.5f, py = (float)y + 0.5f;
    if (R->cam->do_jittering)
    {
        px += orc_rand01(&st) - 0.5f;
        py += orc_rand01(&st) - 0.5f;
    }
    orc_ray ray;
    orc_camera_ray(R->cam, px, py, w, h, &ray);
    v3 o = V(ray.o[0], ray.o[1], ray.o[2]), d = V(ray.d[0], ray.d[1], ray.d[2]);
    hitinfo hi;
    if (trace_main(R, o, d, -1, &hi))
    {
        g->gn = oct_pack(hi.gn);
        g->sn = oct_pack(hi.sn);
        g->mat = R->mat_idx[hi.prim];
        g->pos = hi.p;
        g->prim = hi.prim;
    }
    else
    {
        g->pos = vadd(o, d);
        g->prim = -1;
    }
}

static rgb megakernel_pixel(const rctx* R, int x, int y, uint32_t rn, const gbuf* g)
{
    uint32_t pix = (uint32_t)x + (uint32_t)y * (uint32_t)R->s.width;
    uint32_t st = orc_pixel_seed(pix, R->sample_number, rn, R->s.freeze_random);
    hitinfo hi;
    hi.p = g->pos;
    hi.gn = vnorm(oct_unpack(g->gn));
    hi.sn = vnorm(oct_unpack(g->sn));
    hi.prim = g->prim;
    v3 cam = V(R->cam->position[0], R->cam->position[1], R->cam->position[2]);
    v3 dir = vnorm(vneg(vnorm(vsub(cam, g->pos))));
    int mat = g->mat;
    rgb thr = C1(1.0f), color = C1(0.0f);
    float path_r = 0.0f;
    int missed = 0;
    int found = g->prim != -1;
    /* RayVolumeState::reconstruct_first_hit (RayVolumeState.h:35-59) */
    ovol vs;
    vol_init(&vs);
    if (found)
    {
        const omat* m0 = &R->mats[mat];
        vol_push(&vs, R->mat_idx[g->prim], m0->dielectric_priority);
        if (m0->dispersion_scale > 0.0f && m0->specular_transmission > 0.0f && vs.wavelength == 0.0f)
            vs.wavelength = -(orc_rand01(&st) * (830.0f - 360.0f) + 360.0f);
    }
    for (int bounce = 0; bounce < R->s.nb_bounces + 1; bounce++)
    {
        if (missed)
            break;
        if (bounce > 0)
        {
            v3 o = hi.p;
            int32_t last = hi.prim;
            found = trace_path(R, o, dir, last, &hi, &vs, &st);
            if (found)
                mat = R->mat_idx[hi.prim];
        }
        if (!found)
        {
            missed = 1;
            color = cadd(color, cmul(env_miss(R, dir, bounce), thr));
            continue;
        }
        const omat* m = &R->mats[mat];
        v3 view = vneg(dir);
        if (bounce > 0 || R->s.enable_direct)
        {
            /* estimate_direct_lighting (Lights.h:364-398) */
            rgb direct = C1(0.0f);
            if ((R->ne > 0 || env_on(R)) && R->restir_out && bounce == 0)
                direct = restir_final_shading(R, m, &hi, view, pix, &vs);
            else if (R->ne > 0 && (R->s.direct_light_sampling_strategy == 4 || (R->ris_later && bounce > 0)))
                direct = cadd(C1(0.0f), sample_lights_ris(R, m, &hi, view, path_r, &st, &vs));
            else if (R->ne > 0)
                direct = cadd(C1(0.0f), sample_one_light_mis(R, m, &hi, view, path_r, &st, &vs));
            /* sample_environment_map (Envmap.h:258-276): ReSTIR DI covers
             * the envmap at bounce 0 */
            rgb env = C1(0.0f);
            if (env_on(R) && !(bounce == 0 && R->s.direct_light_sampling_strategy == 5))
                env = sample_envmap_mis(R, m, &hi, view, path_r, &st, &vs);
            rgb total = C1(0.0f);
            if (bounce == 0)
                total = cadd(total, m->emission);
            total = cadd(total, cmul(cadd(cdivs(direct, 1.0f), env), thr));
            color = cadd(color, total);
        }
        /* path_tracing_compute_next_indirect_bounce (PathTracing.h:26-101) */
        bctx c = {view, hi.sn, hi.gn, V(0, 0, 0), NO_INFO, NO_REG, path_r, &vs, 1};
        v3 nd;
        float bpdf;
        rgb f = principled_sample(R, m, &c, &nd, &bpdf, &st);
        /* RayPayload::accumulate_roughness (RayPayload.h:52-89) */
        if (c.info == FROM_DIFFUSE || c.info == FROM_DIFFUSE_TRANSMISSION)
            path_r = 1.0f;
        else if (c.info == FROM_COAT)
            path_r = hmax(m->coat_roughness, path_r);
        else if (c.info == FROM_FIRST_METAL || c.info == FROM_SPECULAR || c.info == FROM_GLASS_REFLECT ||
                 c.info == FROM_GLASS_REFRACT)
            path_r = hmax(m->roughness, path_r);
        else if (c.info == FROM_SECOND_METAL)
            path_r = hmax(m->second_roughness, path_r);
        if (bpdf <= 0.0f)
            break;
        rgb att = cdivs(cscale(f, fabsf(vdot(nd, hi.sn))), bpdf);
        /* do_russian_roulette (RussianRoulette.h:20-56), MAX_THROUGHPUT */
        if (bounce >= R->s.russian_roulette_min_depth && R->s.do_russian_roulette)
        {
            float sp = hmin(cmax(thr), 1.0f);
            if (orc_rand01(&st) > sp)
                break;
            float sc = 1.0f / sp;
            if (R->s.russian_roulette_throughput_clamp > 0.0f)
                sc = hmin(sc, R->s.russian_roulette_throughput_clamp);
            thr = cscale(thr, sc);
        }
        thr = cmul(thr, dispersion_filter(&vs.wavelength, m->dispersion_scale));
        thr = cmul(thr, att);
        thr = C3(hmax(thr.r, 1.0e-5f), hmax(thr.g, 1.0e-5f), hmax(thr.b, 1.0e-5f));
        if (cblack(thr))
            break;
        dir = nd;
    }
    return color;
}

float orc_consistent_tau(float tau_0, uint32_t sample_number)
{
    /* MicrofacetRegularization.h:148-161 */
    float p = powf((float)(int)(sample_number + 1), -1.0f / 6.0f);
    float inner = p * sqrtf(M_FOUR_PI_F * tau_0 - 1.0f) / (M_TWO_PI_F * tau_0 - 1.0f);
    return 1.0f / (2.0f * M_PI_F * (1.0f - cosf(atanf(inner))));
}

/* forward declarations of the frame renderer below */
static int32_t render_sample_via_frame(const orc_render_scene* sc, const orc_octree* tree, const orc_camera* cam,
                                       const orc_render_settings* s, uint32_t sample_number, uint32_t rn_camera,
                                       uint32_t rn_path, float* fb, double* fb64, int32_t* gb_prims, int32_t nthreads);

/* One plain sample (CAMERA_RAYS then PATH_TRACE) on a fresh frame. */
int32_t orc_render_sample(const orc_render_scene* sc, const orc_octree* tree, const orc_camera* cam,
                          const orc_render_settings* s, uint32_t sample_number, uint32_t rn_camera,
                          uint32_t rn_path, float* fb, double* fb64, int32_t* gb_prims, int32_t nthreads)
{
    return render_sample_via_frame(sc, tree, cam, s, sample_number, rn_camera, rn_path, fb, fb64, gb_prims, nthreads);
}

/* ------------------------------------------------------------------ */
/* BSDF entry points for the oracle's own property tests.               */
/* ------------------------------------------------------------------ */
static rctx bsdf_ctx(const orc_luts* luts, float consistent_tau, int32_t regularize)
{
    rctx R;
    memset(&R, 0, sizeof(R));
    R.lut = luts->conductor;
    R.lut_glossy = luts->glossy_dielectric;
    R.lut_sheen = luts->sheen_ltc;
    R.lut_glass = luts->glass;
    R.lut_glass_inv = luts->glass_inv;
    R.lut_thin_glass = luts->thin_glass;
    R.glass_exp = luts->glass_exponent;
    R.consistent_tau = consistent_tau;
    R.s.do_microfacet_regularization = regularize;
    return R;
}

/* A material table (at most 16 materials), the evaluated material and the
 * path's volume state, built by pushing q->pushes (material indices) on an
 * empty stack, then setting the distance and wavelength. */
typedef struct {
    omat mats[16];
    ovol vs;
} bsdf_setup;

static int setup_query(const orc_bsdf_query* q, rctx* R, bsdf_setup* b)
{
    if (q->nmats < 1 || q->nmats > 16 || q->mat < 0 || q->mat >= q->nmats)
        return -1;
    for (int i = 0; i < q->nmats; i++)
        b->mats[i] = unpack_material(q->mats + i);
    R->mats = b->mats;
    R->consistent_tau_refract = q->consistent_tau_refract;
    vol_init(&b->vs);
    for (int i = 0; i < q->npush; i++)
    {
        if (q->pushes[i] < 0 || q->pushes[i] >= q->nmats)
            return -1;
        vol_push(&b->vs, q->pushes[i], b->mats[q->pushes[i]].dielectric_priority);
    }
    b->vs.distance = q->distance;
    b->vs.wavelength = q->wavelength;
    return 0;
}

static void export_state(const ovol* v, float* out)
{
    if (!out)
        return;
    out[0] = (float)v->pos;
    out[1] = (float)v->inside;
    out[2] = (float)v->incident;
    out[3] = (float)v->outgoing;
    out[4] = v->distance;
    out[5] = v->wavelength;
}

int32_t orc_bsdf_eval_q(const orc_bsdf_query* q, const orc_luts* luts, const float* view, const float* sn,
                        const float* gn, const float* to_light, float* out_rgb, float* out_pdf, float* out_state)
{
    rctx R = bsdf_ctx(luts, q->consistent_tau, q->regularize);
    bsdf_setup b;
    if (setup_query(q, &R, &b))
        return -1;
    bctx c = {V(view[0], view[1], view[2]), V(sn[0], sn[1], sn[2]), V(gn[0], gn[1], gn[2]),
              V(to_light[0], to_light[1], to_light[2]), q->info, q->reg_mode, q->path_roughness, &b.vs, q->update};
    rgb f = principled_eval(&R, &b.mats[q->mat], &c, out_pdf);
    out_rgb[0] = f.r; out_rgb[1] = f.g; out_rgb[2] = f.b;
    export_state(&b.vs, out_state);
    return 0;
}

int32_t orc_bsdf_sample_q(const orc_bsdf_query* q, const orc_luts* luts, const float* view, const float* sn,
                          const float* gn, uint32_t* rng_state, float* out_dir, float* out_rgb, float* out_pdf,
                          int32_t* out_info, float* out_state)
{
    rctx R = bsdf_ctx(luts, q->consistent_tau, q->regularize);
    bsdf_setup b;
    if (setup_query(q, &R, &b))
        return -1;
    bctx c = {V(view[0], view[1], view[2]), V(sn[0], sn[1], sn[2]), V(gn[0], gn[1], gn[2]), V(0, 0, 0), NO_INFO,
              q->reg_mode, q->path_roughness, &b.vs, q->update};
    v3 d = V(0, 0, 0);
    rgb f = principled_sample(&R, &b.mats[q->mat], &c, &d, out_pdf, rng_state);
    out_dir[0] = d.x; out_dir[1] = d.y; out_dir[2] = d.z;
    out_rgb[0] = f.r; out_rgb[1] = f.g; out_rgb[2] = f.b;
    if (out_info)
        *out_info = c.info;
    export_state(&b.vs, out_state);
    return 0;
}

float orc_bsdf_pdf_q(const orc_bsdf_query* q, const orc_luts* luts, const float* view, const float* sn,
                     const float* gn, const float* to_light)
{
    rctx R = bsdf_ctx(luts, q->consistent_tau, q->regularize);
    bsdf_setup b;
    if (setup_query(q, &R, &b))
        return -1.0f;
    bctx c = {V(view[0], view[1], view[2]), V(sn[0], sn[1], sn[2]), V(gn[0], gn[1], gn[2]),
              V(to_light[0], to_light[1], to_light[2]), q->info, q->reg_mode, q->path_roughness, &b.vs, 0};
    return principled_pdf(&R, &b.mats[q->mat], &c);
}

/* the single-material forms: the material alone, entered from the air */
static orc_bsdf_query single(const orc_cpu_material* cm, float consistent_tau, int32_t reg_mode, float path_roughness)
{
    static const int32_t push0[1] = {0};
    orc_bsdf_query q;
    memset(&q, 0, sizeof(q));
    q.mats = cm;
    q.nmats = 1;
    q.mat = 0;
    q.pushes = push0;
    q.npush = 1;
    q.consistent_tau = consistent_tau;
    q.consistent_tau_refract = consistent_tau;
    q.regularize = 1;
    q.reg_mode = reg_mode;
    q.path_roughness = path_roughness;
    return q;
}

void orc_bsdf_eval(const orc_cpu_material* cm, const orc_luts* luts, float consistent_tau, int32_t reg_mode,
                   float path_roughness, const float* view, const float* sn, const float* gn, const float* to_light,
                   float* out_rgb, float* out_pdf)
{
    orc_bsdf_query q = single(cm, consistent_tau, reg_mode, path_roughness);
    orc_bsdf_eval_q(&q, luts, view, sn, gn, to_light, out_rgb, out_pdf, NULL);
}

void orc_bsdf_sample(const orc_cpu_material* cm, const orc_luts* luts, float consistent_tau, int32_t reg_mode,
                     float path_roughness, const float* view, const float* sn, const float* gn, uint32_t* rng_state,
                     float* out_dir, float* out_rgb, float* out_pdf)
{
    orc_bsdf_query q = single(cm, consistent_tau, reg_mode, path_roughness);
    orc_bsdf_sample_q(&q, luts, view, sn, gn, rng_state, out_dir, out_rgb, out_pdf, NULL, NULL);
}

float orc_bsdf_pdf(const orc_cpu_material* cm, const orc_luts* luts, float consistent_tau, int32_t reg_mode,
                   float path_roughness, const float* view, const float* sn, const float* gn, const float* to_light)
{
    orc_bsdf_query q = single(cm, consistent_tau, reg_mode, path_roughness);
    return orc_bsdf_pdf_q(&q, luts, view, sn, gn, to_light);
}

/* The nested-dielectrics stack on its own: pushes (mat >= 0, its priority)
 * and pops (mat = -1 - inside flag) applied in order; out[i] receives the
 * state after step i: pos, inside, incident, outgoing, skip, entries 0-3. */
void orc_vol_script(const int32_t* mats, const int32_t* prios, int32_t n, int32_t* out)
{
    ovol v;
    vol_init(&v);
    for (int i = 0; i < n; i++)
    {
        int skip = 0;
        if (mats[i] >= 0)
            skip = vol_push(&v, mats[i], prios[i]);
        else
            vol_pop(&v, -1 - mats[i]);
        int32_t* o = out + 9 * i;
        o[0] = v.pos;
        o[1] = v.inside;
        o[2] = v.incident;
        o[3] = v.outgoing;
        o[4] = skip;
        for (int k = 0; k < 4; k++)
            o[5 + k] = (int32_t)v.e[k];
    }
}

/* glass correction exponent and dispersion helpers (property tests) */
float orc_glass_correction_exponent(const float* table, float roughness, float relative_eta)
{
    return correction_exponent(table, roughness, relative_eta);
}

void orc_thin_film_fresnel(const orc_cpu_material* cm, float ambient_ior, float HoL, float* out)
{
    omat m = unpack_material(cm);
    rgb F = thin_film(&m, ambient_ior, HoL);
    out[0] = F.r; out[1] = F.g; out[2] = F.b;
}

void orc_hue_shift(const float* in, float degrees, float* out)
{
    rgb c = hue_shift(C3(in[0], in[1], in[2]), degrees);
    out[0] = c.r; out[1] = c.g; out[2] = c.b;
}

float orc_dispersion_ior(float abbe, float scale, float ior, float wavelength)
{
    return disp_ior(abbe, scale, ior, wavelength);
}

void orc_wavelength_rgb(float w, float* out)
{
    rgb c = wavelength_rgb(w);
    out[0] = c.r; out[1] = c.g; out[2] = c.b;
}

/* ================================================================== */
/* Frame renderer: persistent G-buffers, power alias table and the     */
/* ReSTIR DI passes, one pixel at a time with inline ray queries.       */
/* ================================================================== */

/* Utils::compute_alias_table (src/Utils/Utils.cpp:63-131): Vose's method on
 * the weights normalised in FP64 to mean 1; FIFO queues of small and large
 * indices. */
void orc_alias_table(const float* input, int32_t n, float total_sum, float* probas, int32_t* alias)
{
    if (n <= 0)
        return;
    double total = total_sum;
    double* sc = (double*)malloc(sizeof(double) * (size_t)n);
    int32_t* small_q = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * n + 1));
    int32_t* large_q = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * n + 1));
    size_t sh = 0, st = 0, lh = 0, lt = 0; /* head / tail of each FIFO */
    for (int32_t i = 0; i < n; i++)
    {
        sc[i] = (double)input[i] / total;
        sc[i] *= (double)n;
        probas[i] = 0.0f;
        alias[i] = 0;
    }
    for (int32_t i = 0; i < n; i++)
    {
        if (sc[i] < 1.0)
            small_q[st++] = i;
        else
            large_q[lt++] = i;
    }
    while (sh < st && lh < lt)
    {
        int32_t s = small_q[sh++], l = large_q[lh++];
        probas[s] = (float)sc[s];
        alias[s] = l;
        sc[l] = (sc[l] + sc[s]) - 1.0;
        if (sc[l] > 1.0)
            large_q[lt++] = l;
        else
            small_q[st++] = l;
    }
    while (lh < lt)
        probas[large_q[lh++]] = 1.0f;
    while (sh < st)
        probas[small_q[sh++]] = 1.0f;
    free(sc);
    free(small_q);
    free(large_q);
}

struct orc_frame {
    rctx R;
    omat* mats;
    float* areas;
    float* alias_probas;
    int32_t* alias_alias;
    orc_restir_settings rs;
    gbuf* cur;
    gbuf* prev;
    ores* res[3];
    int32_t last_output, shade_buffer, temporal_output, odd_frame;
    size_t n;
    /* pixel window [x0, x1) x [y0, y1) that the passes compute (the whole
     * image by default); pixels outside it keep their previous state */
    int32_t win[4];
    /* light presampling: subsets x subset_size presampled lights */
    struct opresampled* presampled;
    size_t presampled_count;
    int32_t fused_output; /* buffer the fused pass (or a spatial pass after it) wrote */
    /* the envmap's packed texels and alias table */
    uint32_t* env_texels;
    float* env_probas;
    int32_t* env_alias;
    /* the camera of the next camera pass (R.cam points here), the camera of
     * the last camera pass and the one before it */
    orc_camera cam, frame_cam, prev_cam;
    int frame_cam_valid;
};

/* ReSTIRDIPresampledLight (ReSTIR/DI/PresampledLight.h:14-30) */
struct opresampled {
    int32_t tri;
    v3 point, normal;
    rgb radiance;
    float pdf;
    uint32_t flags;
};

static int in_window(const orc_frame* f, size_t pix)
{
    const size_t w = (size_t)f->R.s.width;
    const int32_t x = (int32_t)(pix % w), y = (int32_t)(pix / w);
    return x >= f->win[0] && x < f->win[1] && y >= f->win[2] && y < f->win[3];
}

void orc_frame_set_window(orc_frame* f, int32_t x0, int32_t x1, int32_t y0, int32_t y1)
{
    f->win[0] = x0 < 0 ? 0 : x0;
    f->win[1] = x1 > f->R.s.width ? f->R.s.width : x1;
    f->win[2] = y0 < 0 ? 0 : y0;
    f->win[3] = y1 > f->R.s.height ? f->R.s.height : y1;
}

static ores ores_empty(void)
{
    ores r;
    r.point = V(0, 0, 0);
    r.target = 0.0f;
    r.M = 0;
    r.tri = -1;
    r.ucw = 0.0f;
    r.flags = 0u;
    r.wsum = 0.0f;
    return r;
}

static void gbuf_reset(gbuf* g)
{
    g->prim = -1;
    g->pos = V(0, 0, 0);
    g->gn = 0u;
    g->sn = 0u;
    g->mat = -1;
}

static int restir_on(const orc_frame* f) { return f->R.s.direct_light_sampling_strategy == 5; }
static int keep_prev(const orc_frame* f) { return restir_on(f) && f->rs.do_temporal_reuse_pass; }

orc_frame* orc_frame_create(const orc_render_scene* sc, const orc_octree* tree, const orc_camera* cam,
                            const orc_render_settings* s, const orc_restir_settings* rs)
{
    orc_frame* f = (orc_frame*)calloc(1, sizeof(orc_frame));
    f->mats = (omat*)malloc(sizeof(omat) * (size_t)(sc->material_count > 0 ? sc->material_count : 1));
    for (int32_t i = 0; i < sc->material_count; i++)
        f->mats[i] = unpack_material(sc->materials + i);
    f->areas = (float*)malloc(sizeof(float) * (size_t)(sc->triangle_count > 0 ? sc->triangle_count : 1));
    for (int32_t t = 0; t < sc->triangle_count; t++)
    {
        const float* a = sc->vertices + 3 * (int64_t)sc->indices[3 * (int64_t)t];
        const float* b = sc->vertices + 3 * (int64_t)sc->indices[3 * (int64_t)t + 1];
        const float* c = sc->vertices + 3 * (int64_t)sc->indices[3 * (int64_t)t + 2];
        v3 nrm = vcross(V(b[0] - a[0], b[1] - a[1], b[2] - a[2]), V(c[0] - a[0], c[1] - a[1], c[2] - a[2]));
        f->areas[t] = vlen(nrm) * 0.5f;
    }
    /* Emissive power list (GPURenderer.cpp:221-295): area from the float
     * cross product, luminance of the CPU material's emission times its
     * strength and global factor, summed in order in binary32. */
    int32_t E = sc->emissive_count;
    f->alias_probas = (float*)calloc((size_t)(E > 0 ? E : 1), sizeof(float));
    f->alias_alias = (int32_t*)calloc((size_t)(E > 0 ? E : 1), sizeof(int32_t));
    float sum = 0.0f;
    if (E > 0)
    {
        float* power = (float*)malloc(sizeof(float) * (size_t)E);
        for (int32_t i = 0; i < E; i++)
        {
            int32_t t = sc->emissive[i];
            const float* a = sc->vertices + 3 * (int64_t)sc->indices[3 * (int64_t)t];
            const float* b = sc->vertices + 3 * (int64_t)sc->indices[3 * (int64_t)t + 1];
            const float* c = sc->vertices + 3 * (int64_t)sc->indices[3 * (int64_t)t + 2];
            v3 nrm = vcross(V(b[0] - a[0], b[1] - a[1], b[2] - a[2]), V(c[0] - a[0], c[1] - a[1], c[2] - a[2]));
            float area = 0.5f * vlen(nrm);
            const orc_cpu_material* m = sc->materials + sc->material_indices[t];
            float l = 0.3086f * m->emission[0] + 0.6094f * m->emission[1] + 0.0820f * m->emission[2];
            power[i] = l * m->emission_strength * m->global_emissive_factor * area;
            sum += power[i];
        }
        orc_alias_table(power, E, sum, f->alias_probas, f->alias_alias);
        free(power);
    }
    rctx R;
    memset(&R, 0, sizeof(R));
    R.verts = sc->vertices;
    R.idx = sc->indices;
    R.mat_idx = sc->material_indices;
    R.mats = f->mats;
    R.emissive = sc->emissive;
    R.ne = sc->emissive_count;
    R.areas = f->areas;
    if (s->do_normal_mapping && sc->has_vertex_normals && sc->vertex_normals)
    {
        R.has_normals = sc->has_vertex_normals;
        R.vnormals = sc->vertex_normals;
    }
    R.lut = sc->luts.conductor;
    R.lut_glossy = sc->luts.glossy_dielectric;
    R.lut_sheen = sc->luts.sheen_ltc;
    R.lut_glass = sc->luts.glass;
    R.lut_glass_inv = sc->luts.glass_inv;
    R.lut_thin_glass = sc->luts.thin_glass;
    R.glass_exp = sc->luts.glass_exponent;
    R.tree = tree;
    R.cam = NULL;
    R.s = *s;
    R.alias_probas = f->alias_probas;
    R.alias_alias = f->alias_alias;
    R.alias_sum = sum;
    /* WorldSettings; the envmap as RendererEnvmap prepares it
     * (RendererEnvmap.cpp:13-20, :80-97): packed texels, then the alias
     * table over the float texels' luminance (Image.cpp:526-546, :652-672) */
    R.env_type = 1;
    R.env_intensity = 1.0f;
    for (int k = 0; k < 3; k++)
    {
        R.env_to_world[k][k] = 1.0f;
        R.env_to_env[k][k] = 1.0f;
    }
    if (sc->world)
    {
        const orc_world* w = sc->world;
        R.env_type = w->ambient_light_type;
        memcpy(R.env_uniform, w->uniform_light_color, sizeof(R.env_uniform));
        R.env_scale_bg = w->envmap_scale_background_intensity;
        R.env_intensity = w->envmap_intensity;
        memcpy(R.env_to_world, w->envmap_to_world, sizeof(R.env_to_world));
        memcpy(R.env_to_env, w->world_to_envmap, sizeof(R.env_to_env));
        if (w->envmap && w->envmap_width > 0 && w->envmap_height > 0)
        {
            size_t count = (size_t)w->envmap_width * (size_t)w->envmap_height;
            f->env_texels = (uint32_t*)malloc(4 * count);
            f->env_probas = (float*)malloc(4 * count);
            f->env_alias = (int32_t*)malloc(4 * count);
            float* lumv = (float*)malloc(4 * count);
            float esum = 0.0f;
            for (size_t i = 0; i < count; i++)
            {
                const float* px = w->envmap + i * (size_t)w->envmap_channels;
                f->env_texels[i] = orc_rgb9e5_pack(px[0], px[1], px[2]);
                lumv[i] = lum(C3(px[0], px[1], px[2]));
                esum += lumv[i];
            }
            orc_alias_table(lumv, (int32_t)count, esum, f->env_probas, f->env_alias);
            free(lumv);
            R.env_sum = esum;
            R.env_w = w->envmap_width;
            R.env_h = w->envmap_height;
            R.env_texels = f->env_texels;
            R.env_probas = f->env_probas;
            R.env_alias = f->env_alias;
        }
        else if (R.env_type == 2)
            R.env_type = 1; /* set_envmap with an empty image */
    }
    f->R = R;
    f->cam = *cam;
    f->R.cam = &f->cam;
    if (rs)
        f->rs = *rs;
    f->n = (size_t)s->width * (size_t)s->height;
    f->cur = (gbuf*)malloc(sizeof(gbuf) * f->n);
    f->prev = (gbuf*)malloc(sizeof(gbuf) * f->n);
    for (size_t i = 0; i < f->n; i++)
    {
        gbuf_reset(f->cur + i);
        gbuf_reset(f->prev + i);
    }
    for (int k = 0; k < 3; k++)
    {
        f->res[k] = (ores*)malloc(sizeof(ores) * f->n);
        for (size_t i = 0; i < f->n; i++)
            f->res[k][i] = ores_empty();
    }
    f->last_output = 1;
    f->shade_buffer = 1;
    f->temporal_output = 0;
    f->odd_frame = 0;
    f->fused_output = 1;
    orc_frame_set_window(f, 0, s->width, 0, s->height);
    if (rs && rs->do_light_presampling)
    {
        f->presampled_count = (size_t)rs->light_presampling_number_of_subsets * (size_t)rs->light_presampling_subset_size;
        f->presampled = (struct opresampled*)calloc(f->presampled_count, sizeof(struct opresampled));
        for (size_t i = 0; i < f->presampled_count; i++)
            f->presampled[i].tri = -1;
    }
    return f;
}

void orc_frame_free(orc_frame* f)
{
    if (!f)
        return;
    free(f->mats);
    free(f->areas);
    free(f->alias_probas);
    free(f->alias_alias);
    free(f->cur);
    free(f->prev);
    free(f->presampled);
    free(f->env_texels);
    free(f->env_probas);
    free(f->env_alias);
    for (int k = 0; k < 3; k++)
        free(f->res[k]);
    free(f);
}

void orc_frame_env_sample(orc_frame* f, uint32_t* st, float* dir, float* radiance, float* pdf)
{
    v3 d;
    rgb r = env_sample(&f->R, &d, pdf, st);
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
    radiance[0] = r.r; radiance[1] = r.g; radiance[2] = r.b;
}

void orc_frame_env_eval(orc_frame* f, const float* dir, float* radiance, float* pdf)
{
    rgb r = env_eval(&f->R, V(dir[0], dir[1], dir[2]), pdf);
    radiance[0] = r.r; radiance[1] = r.g; radiance[2] = r.b;
}

void orc_frame_env_tables(const orc_frame* f, float* probas, int32_t* alias, float* sum, uint32_t* texels)
{
    size_t count = (size_t)f->R.env_w * (size_t)f->R.env_h;
    if (probas)
        memcpy(probas, f->env_probas, 4 * count);
    if (alias)
        memcpy(alias, f->env_alias, 4 * count);
    if (sum)
        *sum = f->R.env_sum;
    if (texels)
        memcpy(texels, f->env_texels, 4 * count);
}

static void set_threads(int32_t nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* CameraRays (CameraRays.h:83-213): reset at sample 0 (reset_render :20-59),
 * previous-frame copy (:117-124), G-buffer write. */
void orc_frame_set_camera(orc_frame* f, const orc_camera* cam) { f->cam = *cam; }

void orc_frame_set_permutation_bits(orc_frame* f, int32_t bits) { f->rs.permutation_sampling_random_bits = bits; }

void orc_frame_camera_pass(orc_frame* f, uint32_t sample_number, uint32_t rn, int32_t nthreads)
{
    f->R.sample_number = sample_number;
    f->R.consistent_tau = orc_consistent_tau(f->R.s.microfacet_regularization_tau_0, sample_number);
    f->R.consistent_tau_refract = orc_consistent_tau(f->R.s.microfacet_regularization_tau_0, sample_number + 1);
    int kp = keep_prev(f);
    /* prev_camera: the previous frame's camera (the first frame's own) */
    f->prev_cam = f->frame_cam_valid ? f->frame_cam : f->cam;
    f->frame_cam = f->cam;
    f->frame_cam_valid = 1;
    /* need_to_reset: not accumulating, a moving camera keeps the G-buffers
     * and reservoirs (GPURenderer::reset, GPURenderer.cpp:740-761) */
    if (sample_number == 0 && f->R.s.accumulate)
    {
        for (size_t i = 0; i < f->n; i++)
        {
            gbuf_reset(f->cur + i);
            if (kp)
                gbuf_reset(f->prev + i);
        }
        if (restir_on(f))
        {
            for (int k = 0; k < 3; k++)
                for (size_t i = 0; i < f->n; i++)
                    f->res[k][i] = ores_empty();
            f->odd_frame = 0;
        }
    }
    set_threads(nthreads);
    const int w = f->R.s.width;
#pragma omp parallel for schedule(dynamic)
    for (int y = 0; y < f->R.s.height; y++)
        for (int x = 0; x < w; x++)
        {
            size_t pix = (size_t)x + (size_t)y * (size_t)w;
            if (!in_window(f, pix))
                continue;
            if (kp)
                f->prev[pix] = f->cur[pix];
            camera_rays_pixel(&f->R, x, y, rn, f->cur + pix);
        }
}

/* ---- ReSTIR DI helpers ---------------------------------------------- */
typedef struct { int32_t mat, prim; v3 view, sn, gn, p; } osurf;

/* the surface's volume state: its material pushed on an empty stack
 * (RayVolumeState::reconstruct_first_hit, RayVolumeState.h:35-59). A surface
 * without a primary hit (prim -1: a previous-frame pixel whose camera ray
 * missed, with the material left from an older hit) keeps the empty stack,
 * as the reference's early return does (RayVolumeState.h:37-39). Its incident
 * and outgoing media are the air: the reference leaves them at -1
 * (RayVolumeState.h:65) and its BSDF evaluation then reads the IOR of
 * material -1 (Principled.h:1646), an out-of-bounds read. Dispersive glass is
 * refused with ReSTIR DI. */
static ovol surf_vol(const orc_frame* f, const osurf* s)
{
    ovol v;
    vol_init(&v);
    if (s->prim >= 0)
        vol_push(&v, f->R.mat_idx[s->prim], f->R.mats[s->mat].dielectric_priority);
    else
        v.incident = v.outgoing = (int)VOL_AIR;
    return v;
}

static osurf surf_of(const orc_frame* f, const gbuf* g, size_t pix)
{
    osurf s;
    /* a previous-frame surface looks from the previous frame's camera
     * (get_pixel_surface_previous_frame, Surface.h:55-71) */
    const orc_camera* c = g == f->prev ? &f->prev_cam : f->R.cam;
    v3 cam = V(c->position[0], c->position[1], c->position[2]);
    s.mat = g[pix].mat;
    s.prim = g[pix].prim;
    s.p = g[pix].pos;
    /* the view direction comes from the current frame's hit position also for
     * a previous-frame surface (Surface.h:62) */
    s.view = vnorm(vsub(cam, f->cur[pix].pos));
    s.sn = vnorm(oct_unpack(g[pix].sn));
    s.gn = vnorm(oct_unpack(g[pix].gn));
    return s;
}

/* ReSTIR_DI_evaluate_target_function (DI/TargetFunction.h:49-98); with_vis
 * traces the shadow ray (Intersect.h:276-324) when the value is non-zero. */
static float target_fn(const orc_frame* f, const osurf* s, v3 point, int32_t tri_i, uint32_t flags, int with_vis)
{
    const rctx* R = &f->R;
    int env = (flags & RESTIR_ENVMAP) != 0;
    if ((tri_i == -1 && !env) || s->mat < 0)
        return 0.0f;
    float dist;
    v3 dir = sample_dir(R, s->p, point, flags, &dist);
    float cosine = vdot(s->sn, dir);
    if (cosine <= 0.0f)
        return 0.0f;
    ovol vs = surf_vol(f, s);
    bctx c = {s->view, s->sn, s->gn, dir, (int)(flags & RESTIR_LOBE_MASK), NO_REG, 0.0f, &vs, 0};
    float bpdf;
    rgb fc = principled_eval(R, &R->mats[s->mat], &c, &bpdf);
    rgb em;
    float geom = 1.0f;
    if (env)
    {
        float epdf;
        em = env_eval(R, dir, &epdf);
    }
    else
    {
        em = R->mats[R->mat_idx[tri_i]].emission;
        v3 a, b, cc;
        tri(R, tri_i, &a, &b, &cc);
        geom = hmax(0.0f, vdot(vnorm(vcross(vsub(b, a), vsub(cc, a))), vneg(dir)));
        geom /= dist * dist;
    }
    float t = lum(cscale(cscale(cmul(fc, em), cosine), geom));
    if (t == 0.0f)
        return 0.0f;
    if (with_vis)
    {
        orc_hit h = trace1(R, s->p, dir, s->prim, INFINITY);
        int occluded = h.prim >= 0 && h.t < dist - 1.0e-4f;
        t = t * (occluded ? 0.0f : 1.0f);
    }
    return t;
}

static void add_candidate(ores* r, v3 point, int32_t tri_i, float target, uint32_t flags, float w, uint32_t* st)
{
    r->M++;
    r->wsum += w;
    if (orc_rand01(st) < w / r->wsum)
    {
        r->point = point;
        r->tri = tri_i;
        r->target = target;
        r->flags = flags;
    }
}

static int combine_with(ores* r, const ores* o, float mis, float target, float jac, uint32_t* st)
{
    if (o->ucw <= 0.0f)
    {
        r->M += o->M;
        return 0;
    }
    float w = mis * target * o->ucw * jac;
    r->M += o->M;
    r->wsum += w;
    if (orc_rand01(st) < w / r->wsum)
    {
        r->point = o->point;
        r->tri = o->tri;
        r->flags = o->flags;
        r->target = target;
        return 1;
    }
    return 0;
}

static void res_end(ores* r) { r->ucw = r->wsum == 0.0f ? 0.0f : 1.0f / r->target * r->wsum; }

static void res_end_normalized(ores* r, float num, float den)
{
    if (r->wsum == 0.0f || r->wsum < 1.0e-10f || r->wsum > 1.0e10f || den == 0.0f || num == 0.0f)
        r->ucw = 0.0f;
    else
        r->ucw = 1.0f / r->target * r->wsum * num / den;
    if (r->M > 1000000)
        r->M = 1000000;
}

/* ReSTIR_DI_visibility_test_kill_reservoir (ReSTIR/Utils.h:25-67) */
static void visibility_kill(const orc_frame* f, ores* r, v3 p, int32_t prim)
{
    if (r->ucw <= 0.0f && r->wsum <= 0.0f)
        return;
    if (r->flags & RESTIR_UNOCCLUDED)
        return;
    float dist;
    v3 dir = sample_dir(&f->R, p, r->point, r->flags, &dist);
    orc_hit h = trace1(&f->R, p, dir, prim, INFINITY);
    if (h.prim >= 0 && h.t < dist - 1.0e-4f)
        r->ucw = -1.0f;
    else
        r->flags |= RESTIR_UNOCCLUDED;
}

static void end_frame(orc_frame* f, int32_t out)
{
    f->shade_buffer = out;
    f->last_output = out;
    f->odd_frame = !f->odd_frame;
}

/* ReSTIR_DI_InitialCandidates (DI/InitialCandidates.h:180-538), envmap
 * candidate probability 0. */
void orc_frame_restir_initial(orc_frame* f, uint32_t sample_number, uint32_t rn, int32_t nthreads)
{
    f->R.sample_number = sample_number;
    f->R.consistent_tau = orc_consistent_tau(f->R.s.microfacet_regularization_tau_0, sample_number);
    f->R.consistent_tau_refract = orc_consistent_tau(f->R.s.microfacet_regularization_tau_0, sample_number + 1);
    const rctx* R = &f->R;
    const int32_t nl = f->rs.number_of_initial_light_candidates, nb = f->rs.number_of_initial_bsdf_candidates;
    /* InitialCandidates.h:451-459: envmap candidates with an envmap (an
     * intensity of 0 counts as none), only those without emissive triangles */
    const float env_prob = env_on(R) ? (R->ne == 0 ? 1.0f : f->rs.envmap_candidate_probability) : 0.0f;
    if (R->ne > 0 || env_on(R))
    {
        set_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
        for (int64_t pi = 0; pi < (int64_t)f->n; pi++)
        {
            size_t pix = (size_t)pi;
            uint32_t st = orc_pixel_seed((uint32_t)pix, sample_number, rn, R->s.freeze_random);
            if (f->cur[pix].prim == -1 || !in_window(f, pix))
                continue;
            osurf s = surf_of(f, f->cur, pix);
            const omat* m = &R->mats[s.mat];
            ores res = ores_empty();
            for (int32_t i = 0; i < nl; i++)
            {
                float cosine = 0.0f, pdf = 0.0f, dist = 0.0f;
                v3 dir = V(0, 0, 0), point = V(0, 0, 0);
                int32_t tri_i = -1;
                uint32_t sflags = 0u;
                if (f->presampled)
                {
                    /* use_presampled_light_candidate (InitialCandidates.h:33-90) */
                    const int32_t w_ = R->s.width, ts = f->rs.light_presampling_tile_size;
                    const int32_t x = (int32_t)(pix % (size_t)w_), y = (int32_t)(pix / (size_t)w_);
                    const int32_t tx = x / ts, ty = y / ts;
                    const int32_t tile = (tx + ty + 1) * (tx + ty) / 2 + ty; /* Cantor pairing */
                    uint32_t sst = rn * (uint32_t)(tile + 1);
                    int32_t subset = orc_random_index(&sst, f->rs.light_presampling_number_of_subsets);
                    int32_t in_subset = orc_random_index(&st, f->rs.light_presampling_subset_size);
                    const struct opresampled* P =
                        f->presampled + (size_t)subset * (size_t)f->rs.light_presampling_subset_size + (size_t)in_subset;
                    tri_i = P->tri;
                    point = P->point;
                    pdf = P->pdf;
                    sflags = P->flags;
                    dir = sample_dir(R, s.p, point, sflags, &dist);
                    cosine = vdot(s.sn, dir);
                }
                else if (orc_rand01(&st) > env_prob)
                {
                    olight L = sample_emissive(R, &st);
                    tri_i = L.tri;
                    point = L.point;
                    pdf = L.area_pdf;
                    v3 to = vsub(point, s.p);
                    dist = vlen(to);
                    dir = vdivs(to, dist);
                    if (pdf > 0.0f)
                    {
                        cosine = hmax(0.0f, vdot(s.sn, dir));
                        pdf *= (1.0f - env_prob);
                    }
                }
                else if (env_prob > 0.0f) /* a draw of 0 with no envmap: no candidate */
                {
                    /* the envmap sample branch of sample_fresh_light_candidate
                     * (InitialCandidates.h:146-175), stored in envmap space */
                    v3 edir;
                    env_sample(R, &edir, &pdf, &st);
                    cosine = hmax(0.0f, vdot(edir, s.sn));
                    pdf *= env_prob;
                    point = env_rot(R->env_to_env, edir);
                    sflags = RESTIR_ENVMAP;
                    dir = sample_dir(R, s.p, point, sflags, &dist);
                }
                float w = 0.0f, sample_target = 0.0f;
                if (cosine > 0.0f && pdf > 0.0f)
                {
                    ovol vs = surf_vol(f, &s);
                    bctx c = {s.view, s.sn, s.gn, dir, NO_INFO, NO_REG, 0.0f, &vs, 0};
                    float bsdf_pdf;
                    principled_eval(R, m, &c, &bsdf_pdf);
                    float target = target_fn(f, &s, point, tri_i, sflags, 0);
                    if (!(bsdf_pdf <= 0.0f))
                    {
                        float lpdf = pdf; /* an envmap pdf is in solid angle already */
                        if (!(sflags & RESTIR_ENVMAP))
                        {
                            v3 a, b, cc;
                            tri(R, tri_i, &a, &b, &cc);
                            v3 ln = vcross(vsub(b, a), vsub(cc, a));
                            float nl_len = vlen(ln);
                            ln = vdivs(ln, nl_len);
                            float cl = hmax(0.0f, vdot(ln, vneg(dir)));
                            lpdf = cl < 1.0e-8f ? 0.0f : (pdf / (1.0f - env_prob)) * (dist * dist) / cl;
                            lpdf *= (1.0f - env_prob);
                        }
                        float mis = lpdf == 0.0f ? 0.0f : lpdf / ((float)nl * lpdf + (float)nb * bsdf_pdf);
                        w = mis * target / pdf;
                        sample_target = target;
                    }
                }
                add_candidate(&res, point, tri_i, sample_target, sflags, w, &st);
            }
            for (int32_t i = 0; i < nb; i++)
            {
                ovol vs = surf_vol(f, &s);
                bctx c = {s.view, s.sn, s.gn, V(0, 0, 0), NO_INFO, NO_REG, 0.0f, &vs, 0};
                v3 dir;
                float bpdf;
                principled_sample(R, m, &c, &dir, &bpdf, &st);
                if (!(bpdf > 0.0f))
                    continue;
                orc_hit h = trace1(R, s.p, dir, s.prim, INFINITY);
                if (!(h.prim >= 0 && h.t < 1.0e35f - 1.0e-4f))
                {
                    /* an envmap hit becomes an envmap sample
                     * (InitialCandidates.h:396-436), refractions excluded */
                    if (env_on(R) && vdot(s.sn, dir) > 0.0f)
                    {
                        float epdf;
                        env_eval(R, dir, &epdf);
                        v3 point = env_rot(R->env_to_env, dir);
                        uint32_t flags = (RESTIR_UNOCCLUDED | RESTIR_ENVMAP | (uint32_t)c.info) & 0xffu;
                        float target = target_fn(f, &s, point, -1, flags, 0);
                        epdf *= env_prob;
                        float mis = bpdf == 0.0f ? 0.0f : bpdf / ((float)nb * bpdf + (float)nl * epdf);
                        float w = mis * target / bpdf;
                        add_candidate(&res, point, -1, target, flags, w, &st);
                    }
                    continue;
                }
                rgb em = R->mats[R->mat_idx[h.prim]].emission;
                if (cblack(em))
                    continue;
                v3 point = vadd(s.p, vscale(dir, h.t));
                uint32_t flags = (RESTIR_UNOCCLUDED | (uint32_t)c.info) & 0xffu;
                float target = target_fn(f, &s, point, h.prim, flags, 0);
                v3 a, b, cc;
                tri(R, h.prim, &a, &b, &cc);
                v3 hgn = vnorm(vnorm(vcross(vsub(b, a), vsub(cc, a))));
                float cl = hmax(0.0f, vdot(hgn, vneg(dir)));
                float lpdf = 0.0f;
                if (!(vdot(dir, s.sn) < 0.0f))
                {
                    /* with presampling, the presampling strategy's pdf:
                     * power (InitialCandidates.h:360-366) */
                    float pa = emissive_hit_area_pdf_with(R, f->presampled ? 1 : R->s.direct_light_sampling_base_strategy,
                                                          R->areas[h.prim], em);
                    lpdf = cl < 1.0e-8f ? 0.0f : pa * (h.t * h.t) / cl;
                }
                lpdf *= (1.0f - env_prob);
                float mis = bpdf == 0.0f ? 0.0f : bpdf / ((float)nb * bpdf + (float)nl * lpdf);
                float pdf_area = bpdf;
                pdf_area /= (h.t * h.t);
                pdf_area *= cl;
                float w = mis * target / pdf_area;
                if (f->rs.guard_backfacing_bsdf_candidates && !(pdf_area > 0.0f))
                    w = 0.0f;
                add_candidate(&res, point, h.prim, target, flags, w, &st);
            }
            res_end(&res);
            res.M = 1;
            visibility_kill(f, &res, s.p, s.prim);
            f->res[0][pix] = res;
        }
    }
    if (!f->rs.do_temporal_reuse_pass && !f->rs.do_spatial_reuse_pass && !f->rs.do_fused_spatiotemporal)
        end_frame(f, 0);
}

/* ReSTIR_DI_TemporalReuse (DI/TemporalReuse.h:50-296), accumulation mode
 * (the temporal neighbour is the pixel itself, UtilsTemporal.h:37-41), with
 * the pairwise MIS weights (TemporalMISWeight.h:164-254). */
/* ReSTIR_DI_LightsPresampling (DI/LightsPresampling.h:74-112): one light
 * per slot, seeded like a pixel by its slot index; the envmap with the
 * envmap candidate probability (presample_envmap, :22-35, direction in
 * envmap space), else an emissive triangle by power
 * (ReSTIR_DI_LightPresamplingStrategy, ReSTIRDIOptions.h:162). */
void orc_frame_restir_presample(orc_frame* f, uint32_t sample_number, uint32_t rn, int32_t nthreads)
{
    if (!f->presampled || (f->R.ne == 0 && !env_on(&f->R)))
        return;
    const rctx* R = &f->R;
    const float env_prob = env_on(R) ? (R->ne == 0 ? 1.0f : f->rs.envmap_candidate_probability) : 0.0f;
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t x = 0; x < (int64_t)f->presampled_count; x++)
    {
        uint32_t st = orc_pixel_seed((uint32_t)x, sample_number, rn, R->s.freeze_random);
        struct opresampled P;
        memset(&P, 0, sizeof(P));
        P.tri = -1;
        if (orc_rand01(&st) < env_prob)
        {
            v3 d;
            P.radiance = env_sample(R, &d, &P.pdf, &st);
            P.point = env_rot(R->env_to_env, d);
            P.pdf *= env_prob;
            P.flags = RESTIR_ENVMAP;
        }
        else
        {
            olight L = sample_emissive_with(R, 1, &st);
            if (L.area_pdf > 0.0f)
            {
                P.point = L.point;
                P.normal = L.normal;
                P.tri = L.tri;
                P.pdf = L.area_pdf * (1.0f - env_prob);
                P.radiance = L.emission;
            }
        }
        f->presampled[x] = P;
    }
}

static int similar_prev(const orc_frame* f, size_t np, size_t cp, v3 cpoint, v3 cnormal);

/* float -> pixel coordinate, out-of-range values to a coordinate far outside
 * any viewport (the reference's int conversion of them is undefined) */
static int pixel_coord(float v) { return (v > -1.0e9f && v < 1.0e9f) ? (int)v : -1000000000; }

/* find_temporal_neighbor_index (UtilsTemporal.h:37-99): the pixel itself
 * when accumulating; else the primary hit reprojected by the previous
 * frame's view-projection (matrix_X_point, Math.h:617-633), then up to
 * max_neighbor_search_count random pixels within neighbor_search_radius
 * (x offset drawn first), the first that passes the similarity heuristics
 * against the previous G-buffer. Returns the index (-1: none) and the
 * rounded reprojected pixel in *px, *py. */
static int64_t temporal_neighbor(const orc_frame* f, size_t pix, v3 cnormal, uint32_t* st, int* px, int* py)
{
    const int W = f->R.s.width, H = f->R.s.height;
    if (f->R.s.accumulate)
    {
        *px = (int)(pix % (size_t)W);
        *py = (int)(pix / (size_t)W);
        return (int64_t)pix;
    }
    const v3 p = f->cur[pix].pos;
    const float(*m)[4] = f->prev_cam.view_projection;
    float xt = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
    float yt = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
    float wt = m[3][0] * p.x + m[3][1] * p.y + m[3][2] * p.z + m[3][3];
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
    int64_t idx = -1;
    for (int i = 0; i < f->rs.max_neighbor_search_count + 1; i++)
    {
        float ox = 0.0f, oy = 0.0f;
        if (i > 0)
        {
            float rx = orc_rand01(st) - 0.5f;
            float ry = orc_rand01(st) - 0.5f;
            ox = rx * (float)f->rs.neighbor_search_radius;
            oy = ry * (float)f->rs.neighbor_search_radius;
        }
        int x = pixel_coord(roundf(fx + ox)), y = pixel_coord(roundf(fy + oy));
        if (f->rs.use_permutation_sampling && i == 0)
        {
            /* apply_permutation_sampling (UtilsTemporal.h:11-22) */
            int bits = f->rs.permutation_sampling_random_bits;
            int dx = bits & 3, dy = (bits >> 2) & 3;
            x += dx;
            y += dy;
            x ^= 3;
            y ^= 3;
            x -= dx;
            y -= dy;
        }
        if (x < 0 || x >= W || y < 0 || y >= H)
            continue;
        idx = (int64_t)x + (int64_t)y * W;
        if (similar_prev(f, (size_t)idx, pix, p, cnormal))
            break;
        idx = -1;
    }
    *px = pixel_coord(roundf(fx));
    *py = pixel_coord(roundf(fy));
    return idx;
}

void orc_frame_temporal_neighbors(orc_frame* f, uint32_t sample_number, uint32_t rn, int32_t* out)
{
    for (size_t pix = 0; pix < f->n; pix++)
    {
        int32_t* o = out + 3 * pix;
        o[0] = -2;
        o[1] = o[2] = 0;
        if (f->cur[pix].prim == -1)
            continue;
        uint32_t st = orc_pixel_seed((uint32_t)pix, sample_number, rn, f->R.s.freeze_random);
        int x, y;
        o[0] = (int32_t)temporal_neighbor(f, pix, vnorm(oct_unpack(f->cur[pix].sn)), &st, &x, &y);
        o[1] = x;
        o[2] = y;
    }
}

void orc_frame_restir_temporal(orc_frame* f, uint32_t sample_number, uint32_t rn, int32_t nthreads)
{
    f->R.sample_number = sample_number;
    f->R.consistent_tau = orc_consistent_tau(f->R.s.microfacet_regularization_tau_0, sample_number);
    f->R.consistent_tau_refract = orc_consistent_tau(f->R.s.microfacet_regularization_tau_0, sample_number + 1);
    const int32_t out_i = f->rs.do_spatial_reuse_pass ? 0 : (f->odd_frame ? 1 : 2);
    ores* in = f->res[f->last_output];
    ores* init = f->res[0];
    ores* out = f->res[out_i];
    const int cw = f->rs.use_confidence_weights != 0;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
    for (int64_t pi = 0; pi < (int64_t)f->n; pi++)
    {
        size_t pix = (size_t)pi;
        if (f->cur[pix].prim == -1 || !in_window(f, pix))
            continue;
        uint32_t st = orc_pixel_seed((uint32_t)pix, sample_number, rn, f->R.s.freeze_random);
        ores ini = init[pix];
        int tx, ty;
        v3 csn = vnorm(oct_unpack(f->cur[pix].sn));
        int64_t tix = temporal_neighbor(f, pix, csn, &st, &tx, &ty);
        if (tix < 0 || f->R.s.freeze_random || in[tix].M == 0)
        {
            out[pix] = ini;
            continue;
        }
        ores tn = in[tix];
        osurf prev = surf_of(f, f->prev, (size_t)tix);
        if (prev.mat >= 0)
        {
            rgb e = f->R.mats[prev.mat].emission;
            const float nz = 1.0e-10f;
            if (!(e.r < nz && e.r > -nz) || !(e.g < nz && e.g > -nz) || !(e.b < nz && e.b > -nz))
            {
                out[pix] = ini;
                continue;
            }
        }
        osurf center = surf_of(f, f->cur, pix);
        ores o = ores_empty();
        float tf_center = 0.0f;
        if (tn.ucw > 0.0f)
            tf_center = target_fn(f, &center, tn.point, tn.tri, tn.flags, 1);
        float tn_M = cw ? (float)tn.M : 1.0f;
        float center_M = cw ? (float)ini.M : 1.0f;
        float conf_sum = cw ? (float)tn.M : 1.0f;
        float nume = tn.target * tn_M;
        float denom = tn.target * conf_sum + tf_center * center_M;
        float mi = denom == 0.0f ? 0.0f : nume / denom;
        float tf_csn = target_fn(f, &prev, ini.point, ini.tri, ini.flags, 1);
        float tf_csc = ini.target;
        float nume_mc = tf_csc * center_M;
        float denom_mc = tf_csn * conf_sum + tf_csc * center_M;
        float conf_mult = cw ? tn_M / conf_sum : 1.0f;
        float mc = 0.0f;
        if (denom_mc != 0.0f)
            mc += nume_mc / denom_mc * conf_mult;
        if (combine_with(&o, &tn, mi, tf_center, 1.0f, &st))
            o.flags |= RESTIR_UNOCCLUDED;
        float mis_init = mc == 0.0f ? 1.0f : mc;
        if (combine_with(&o, &ini, mis_init, ini.target, 1.0f, &st))
            o.flags |= RESTIR_UNOCCLUDED;
        res_end_normalized(&o, 1.0f, 1.0f);
        if (f->rs.m_cap > 0 && o.M > f->rs.m_cap)
            o.M = f->rs.m_cap;
        out[pix] = o;
    }
    f->temporal_output = out_i;
    if (!f->rs.do_spatial_reuse_pass)
        end_frame(f, out_i);
}

/* get_spatial_neighbor_pixel_index (UtilsSpatial.h:177-259) with
 * sample_in_disk_uv (Sampling.h:50-57); the u draw is evaluated first. */
static int64_t spatial_neighbor(const orc_frame* f, int cx, int cy, uint32_t* nst)
{
    float u = orc_rand01(nst);
    float v = orc_rand01(nst);
    float rad = (float)f->rs.spatial_reuse_radius * sqrtf(v);
    float sn_, cs_;
    orc_sincos(M_TWO_PI_F * u, &sn_, &cs_);
    float ox = rad * cs_, oy = rad * sn_;
    int nx = cx + (int)roundf(ox), ny = cy + (int)roundf(oy);
    if (nx < 0 || nx >= f->R.s.width || ny < 0 || ny >= f->R.s.height)
        return -1;
    return (int64_t)nx + (int64_t)ny * f->R.s.width;
}

/* check_neighbor_similarity_heuristics (NeighborSimilarity.h:46-107) */
static int similar(const orc_frame* f, size_t np, size_t cp, v3 cpoint, v3 cnormal)
{
    if (np == cp)
        return 1;
    if (f->cur[np].prim == -1)
        return 0;
    v3 npt = f->cur[np].pos;
    v3 nn = f->rs.reject_using_geometric_normals ? vnorm(oct_unpack(f->cur[np].gn)) : vnorm(oct_unpack(f->cur[np].sn));
    int plane = 1, normal = 1;
    if (f->rs.use_plane_distance_heuristic)
        plane = fabsf(vdot(vsub(npt, cpoint), cnormal)) < f->rs.plane_distance_threshold;
    if (f->rs.use_normal_similarity_heuristic)
        normal = vdot(cnormal, nn) > f->rs.normal_similarity_angle_precomp;
    return plane && normal;
}

/* ReSTIR_DI_SpatialReuse (DI/SpatialReuse.h:43-322) with the pairwise MIS
 * weights (SpatialMISWeight.h:118-234) and confidence weights. */
void orc_frame_restir_spatial(orc_frame* f, int32_t pass_index, uint32_t sample_number, uint32_t rn, int32_t nthreads)
{
    f->R.sample_number = sample_number;
    f->R.consistent_tau = orc_consistent_tau(f->R.s.microfacet_regularization_tau_0, sample_number);
    f->R.consistent_tau_refract = orc_consistent_tau(f->R.s.microfacet_regularization_tau_0, sample_number + 1);
    int32_t in_i, out_i;
    if (f->rs.do_fused_spatiotemporal)
    {
        /* configure_spatial_pass_for_fused_spatiotemporal
         * (ReSTIRDIRenderPass.cpp:573-604): after the fused pass, each pass
         * reads the previous output and writes the other buffer */
        in_i = f->fused_output;
        out_i = in_i == 1 ? 2 : 1;
    }
    else if (pass_index == 0)
    {
        in_i = f->rs.do_temporal_reuse_pass ? f->temporal_output : 0;
        out_i = 1;
    }
    else if ((pass_index & 1) == 0)
    {
        in_i = 2;
        out_i = 1;
    }
    else
    {
        in_i = 1;
        out_i = 2;
    }
    int last = pass_index == f->rs.spatial_number_of_passes - 1;
    int kill_after = f->rs.do_temporal_reuse_pass || !last;
    ores* in = f->res[in_i];
    ores* out = f->res[out_i];
    const int32_t K = f->rs.spatial_reuse_neighbor_count;
    const int cw = f->rs.use_confidence_weights != 0;
    const int w = f->R.s.width;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
    for (int64_t pi = 0; pi < (int64_t)f->n; pi++)
    {
        size_t pix = (size_t)pi;
        if (f->cur[pix].prim == -1 || !in_window(f, pix))
            continue;
        int cx = (int)(pix % (size_t)w), cy = (int)(pix / (size_t)w);
        uint32_t st = orc_pixel_seed((uint32_t)pix, sample_number, rn, f->R.s.freeze_random);
        ores center = in[pix];
        osurf cs = surf_of(f, f->cur, pix);
        uint32_t nseed = orc_xorshift32(&st);
        v3 cnormal = f->rs.reject_using_geometric_normals ? cs.gn : cs.sn;
        int32_t valid_count = 0, valid_M = 0;
        uint32_t cache = 0u;
        uint32_t nst = nseed;
        for (int32_t i = 0; i < K; i++)
        {
            int64_t np = spatial_neighbor(f, cx, cy, &nst);
            if (np < 0)
                continue;
            if (!similar(f, (size_t)np, pix, cs.p, cnormal))
                continue;
            valid_M += in[np].M;
            valid_count++;
            cache |= 1u << i;
        }
        float center_M = cw ? (float)center.M : 1.0f;
        float conf_sum = cw ? (float)valid_M : 1.0f;
        float vdt = cw ? 1.0f : (float)valid_count;
        int update_mc = center.M > 0 && center.ucw > 0.0f;
        float tf_csc = center.target;
        float nume_mc = tf_csc / vdt * center_M;
        float mc = 0.0f;
        ores o = ores_empty();
        nst = nseed;
        int32_t start = valid_M == 0 ? K : 0;
        for (int32_t i = start; i < K + 1; i++)
        {
            if (i < K && !(cache & (1u << i)))
            {
                orc_xorshift32(&nst);
                orc_xorshift32(&nst);
                continue;
            }
            int64_t np = i == K ? (int64_t)pix : spatial_neighbor(f, cx, cy, &nst);
            if (np < 0)
                continue;
            ores nr = in[np];
            float tf_center = 0.0f;
            if (nr.ucw > 0.0f)
                tf_center = i == K ? nr.target : target_fn(f, &cs, nr.point, nr.tri, nr.flags, 0);
            float mis;
            if (i < K)
            {
                float res_M = cw ? (float)nr.M : 1.0f;
                float nume = nr.target * res_M;
                float denom = nr.target * conf_sum + tf_center / vdt * center_M;
                float mi = denom == 0.0f ? 0.0f : nume / denom;
                if (update_mc)
                {
                    osurf ns = surf_of(f, f->cur, (size_t)np);
                    float tf_csn = target_fn(f, &ns, center.point, center.tri, center.flags, 1);
                    float denom_mc = tf_csn * conf_sum + nume_mc;
                    float conf_mult = cw ? (conf_sum == 0.0f ? 0.0f : res_M / conf_sum) : 1.0f;
                    if (denom_mc != 0.0f)
                        mc += nume_mc / denom_mc / vdt * conf_mult;
                }
                mis = mi / vdt;
            }
            else
                mis = mc == 0.0f ? 1.0f : mc;
            if (combine_with(&o, &nr, mis, tf_center, 1.0f, &st))
            {
                if (i == K)
                    o.flags |= nr.flags & RESTIR_UNOCCLUDED;
                else
                    o.flags &= ~RESTIR_UNOCCLUDED;
            }
        }
        res_end_normalized(&o, 1.0f, 1.0f);
        if (kill_after)
            visibility_kill(f, &o, cs.p, cs.prim);
        if (f->rs.m_cap > 0 && o.M > f->rs.m_cap)
            o.M = f->rs.m_cap;
        out[pix] = o;
    }
    if (f->rs.do_fused_spatiotemporal)
        f->fused_output = out_i;
    if (last)
        end_frame(f, out_i);
}

/* check_neighbor_similarity_heuristics(..., previous_frame = 1)
 * (NeighborSimilarity.h:52-107): the neighbour's primary hit and position
 * from the previous frame's G-buffer, its normal from the current one */
static int similar_prev(const orc_frame* f, size_t np, size_t cp, v3 cpoint, v3 cnormal)
{
    if (np == cp)
        return 1;
    if (f->prev[np].prim == -1)
        return 0;
    v3 npt = f->prev[np].pos;
    v3 nn = f->rs.reject_using_geometric_normals ? vnorm(oct_unpack(f->cur[np].gn)) : vnorm(oct_unpack(f->cur[np].sn));
    int plane = 1, normal = 1;
    if (f->rs.use_plane_distance_heuristic)
        plane = fabsf(vdot(vsub(npt, cpoint), cnormal)) < f->rs.plane_distance_threshold;
    if (f->rs.use_normal_similarity_heuristic)
        normal = vdot(cnormal, nn) > f->rs.normal_similarity_angle_precomp;
    return plane && normal;
}

/* ReSTIR_DI_SpatiotemporalReuse (DI/FusedSpatiotemporalReuse.h:152-566) with
 * a static camera: the temporal neighbour is the pixel itself
 * (UtilsTemporal.h:37-41), its reservoir and the spatial neighbours' come from
 * last frame's output, neighbour surfaces and similarity tests from the
 * previous frame's G-buffer (use_prev_frame_g_buffer()), pairwise MIS
 * (SpatiotemporalMISWeight.h:139-224): the temporal neighbour resampled
 * first, the spatial neighbours around it, the initial candidates last.
 * Seeded with wang_hash(((pix + 1) * (sample + 1)) ^ random) (:173). */
void orc_frame_restir_spatiotemporal(orc_frame* f, uint32_t sample_number, uint32_t rn, int32_t nthreads)
{
    f->R.sample_number = sample_number;
    f->R.consistent_tau = orc_consistent_tau(f->R.s.microfacet_regularization_tau_0, sample_number);
    f->R.consistent_tau_refract = orc_consistent_tau(f->R.s.microfacet_regularization_tau_0, sample_number + 1);
    const int32_t in_i = f->last_output, out_i = in_i == 1 ? 2 : 1;
    ores* in = f->res[in_i];
    ores* init_buf = f->res[0];
    ores* out = f->res[out_i];
    const int32_t K = f->rs.spatial_reuse_neighbor_count;
    const int cw = f->rs.use_confidence_weights != 0;
    const int kill_after = f->rs.do_temporal_reuse_pass || f->rs.spatial_number_of_passes - 1 != 0;
    const int w = f->R.s.width;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
    for (int64_t pi = 0; pi < (int64_t)f->n; pi++)
    {
        size_t pix = (size_t)pi;
        if (f->cur[pix].prim == -1 || !in_window(f, pix))
            continue;
        int cx = (int)(pix % (size_t)w), cy = (int)(pix / (size_t)w);
        uint32_t st = f->R.s.freeze_random ? orc_wang_hash((uint32_t)pix + 1u)
                                           : orc_wang_hash((((uint32_t)pix + 1u) * (sample_number + 1u)) ^ rn);
        uint32_t nseed = orc_xorshift32(&st);
        osurf cs = surf_of(f, f->cur, pix);
        /* load_spatiotemporal_neighbor_data (:71-101); the spatial
         * neighbours are drawn around the reprojected pixel */
        int tx, ty;
        int64_t tix = temporal_neighbor(f, pix, cs.sn, &st, &tx, &ty);
        ores tn = (tix < 0 || f->R.s.freeze_random) ? ores_empty() : in[tix];
        cx = tx;
        cy = ty;
        v3 cnormal = f->rs.reject_using_geometric_normals ? cs.gn : cs.sn;
        int32_t valid_count = 0, valid_M = 0;
        uint32_t cache = 0u;
        uint32_t nst = nseed;
        for (int32_t i = 0; i < K; i++)
        {
            int64_t np = spatial_neighbor(f, cx, cy, &nst);
            if (np < 0)
                continue;
            if (!similar_prev(f, (size_t)np, pix, cs.p, cnormal))
                continue;
            valid_M += in[np].M;
            valid_count++;
            cache |= 1u << i;
        }
        if (tn.M > 0)
        {
            valid_count++;
            valid_M += tn.M;
        }
        ores init = init_buf[pix];
        float center_M = cw ? (float)init.M : 1.0f;
        float conf_sum = cw ? (float)valid_M : 1.0f;
        float vdt = cw ? 1.0f : (float)valid_count;
        int update_mc = init.M > 0 && init.ucw > 0.0f;
        float nume_mc = init.target / vdt * center_M;
        float mc = 0.0f;
        ores o = ores_empty();
        if (tn.M > 0)
        {
            float tf_center = tn.ucw > 0.0f ? target_fn(f, &cs, tn.point, tn.tri, tn.flags, 1) : 0.0f;
            float tn_M = cw ? (float)tn.M : 1.0f;
            float nume = tn.target * tn_M;
            float denom = tn.target * conf_sum + tf_center / vdt * center_M;
            float mi = denom == 0.0f ? 0.0f : nume / denom;
            if (update_mc)
            {
                osurf ts = surf_of(f, f->prev, (size_t)tix);
                float tf_csn = target_fn(f, &ts, init.point, init.tri, init.flags, 1);
                float denom_mc = tf_csn * conf_sum + nume_mc;
                float conf_mult = cw ? tn_M / conf_sum : 1.0f;
                if (denom_mc != 0.0f)
                    mc += nume_mc / denom_mc / vdt * conf_mult;
            }
            if (combine_with(&o, &tn, mi / vdt, tf_center, 1.0f, &st))
                o.flags |= RESTIR_UNOCCLUDED;
        }
        nst = nseed;
        int32_t start = valid_M == 0 ? K : 0;
        for (int32_t i = start; i < K + 1; i++)
        {
            if (i < K && !(cache & (1u << i)))
            {
                orc_xorshift32(&nst);
                orc_xorshift32(&nst);
                continue;
            }
            int64_t np = i == K ? (int64_t)pix : spatial_neighbor(f, cx, cy, &nst);
            if (np < 0)
                continue;
            ores nr = i == K ? init : in[np];
            float tf_center = 0.0f;
            if (nr.ucw > 0.0f)
                tf_center = i == K ? nr.target : target_fn(f, &cs, nr.point, nr.tri, nr.flags, 0);
            float mis;
            if (i < K)
            {
                float res_M = cw ? (float)nr.M : 1.0f;
                float nume = nr.target * res_M;
                float denom = nr.target * conf_sum + tf_center / vdt * center_M;
                float mi = denom == 0.0f ? 0.0f : nume / denom;
                if (update_mc)
                {
                    osurf ns = surf_of(f, f->prev, (size_t)np);
                    float tf_csn = target_fn(f, &ns, init.point, init.tri, init.flags, 1);
                    float denom_mc = tf_csn * conf_sum + nume_mc;
                    float conf_mult = cw ? res_M / conf_sum : 1.0f;
                    if (denom_mc != 0.0f)
                        mc += nume_mc / denom_mc / vdt * conf_mult;
                }
                mis = mi / vdt;
            }
            else
                mis = mc == 0.0f ? 1.0f : mc;
            if (combine_with(&o, &nr, mis, tf_center, 1.0f, &st))
            {
                if (i == K)
                    o.flags |= nr.flags & RESTIR_UNOCCLUDED;
                else
                    o.flags &= ~RESTIR_UNOCCLUDED;
            }
        }
        res_end_normalized(&o, 1.0f, 1.0f);
        if (kill_after)
            visibility_kill(f, &o, cs.p, cs.prim);
        if (f->rs.m_cap > 0 && o.M > f->rs.m_cap)
            o.M = f->rs.m_cap;
        out[pix] = o;
    }
    f->fused_output = out_i;
    if (f->rs.spatial_number_of_passes == 1)
        end_frame(f, out_i);
}

int32_t orc_frame_path_trace(orc_frame* f, uint32_t sample_number, uint32_t rn, float* fb, double* fb64,
                             int32_t nthreads)
{
    f->R.sample_number = sample_number;
    f->R.consistent_tau = orc_consistent_tau(f->R.s.microfacet_regularization_tau_0, sample_number);
    f->R.consistent_tau_refract = orc_consistent_tau(f->R.s.microfacet_regularization_tau_0, sample_number + 1);
    f->R.restir_out = restir_on(f) ? f->res[f->shade_buffer] : NULL;
    f->R.final_shading_visibility = f->rs.do_final_shading_visibility;
    f->R.ris_later = restir_on(f) && f->rs.later_bounces_sampling_strategy == 3;
    int32_t rejected = 0;
    const int w = f->R.s.width;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic) reduction(+ : rejected)
    for (int y = 0; y < f->R.s.height; y++)
        for (int x = 0; x < w; x++)
        {
            size_t pix = (size_t)x + (size_t)y * (size_t)w;
            if (!in_window(f, pix))
                continue;
            const gbuf* g = f->cur + pix;
            /* a camera ray that missed gathers the ambient light at bounce 0 */
            rgb c = megakernel_pixel(&f->R, x, y, rn, g);
            int bad = 0;
            float comp[3] = {c.r, c.g, c.b};
            for (int k = 0; k < 3; k++)
                bad |= comp[k] < 0.0f || isnan(comp[k]) || isinf(comp[k]);
            if (bad)
            {
                rejected++;
                c = C1(0.0f);
                if (sample_number != 0)
                    continue;
            }
            float* fp = fb + 3 * pix;
            if (sample_number == 0)
            {
                fp[0] = c.r; fp[1] = c.g; fp[2] = c.b;
            }
            else
            {
                fp[0] += c.r; fp[1] += c.g; fp[2] += c.b;
            }
            if (fb64)
            {
                double* d = fb64 + 3 * pix;
                if (sample_number == 0)
                {
                    d[0] = c.r; d[1] = c.g; d[2] = c.b;
                }
                else
                {
                    d[0] += (double)c.r; d[1] += (double)c.g; d[2] += (double)c.b;
                }
            }
        }
    return rejected;
}

void orc_frame_read_reservoirs(const orc_frame* f, int32_t which, void* out)
{
    const ores* r = f->res[which == 3 ? f->shade_buffer : which];
    uint32_t* o = (uint32_t*)out;
    for (size_t i = 0; i < f->n; i++)
    {
        float fl[4] = {r[i].point.x, r[i].point.y, r[i].point.z, r[i].target};
        memcpy(o + 8 * i, fl, 16);
        int32_t iv[2] = {r[i].M, r[i].tri};
        memcpy(o + 8 * i + 4, iv, 8);
        memcpy(o + 8 * i + 6, &r[i].ucw, 4);
        memcpy(o + 8 * i + 7, &r[i].flags, 4);
    }
}

void orc_frame_gbuffer_prims(const orc_frame* f, int32_t* out)
{
    for (size_t i = 0; i < f->n; i++)
        out[i] = f->cur[i].prim;
}

void orc_frame_alias_table(const orc_frame* f, float* probas, int32_t* alias, float* sum)
{
    int32_t E = f->R.ne;
    if (probas)
        memcpy(probas, f->alias_probas, sizeof(float) * (size_t)E);
    if (alias)
        memcpy(alias, f->alias_alias, sizeof(int32_t) * (size_t)E);
    if (sum)
        *sum = f->R.alias_sum;
}

static int32_t render_sample_via_frame(const orc_render_scene* sc, const orc_octree* tree, const orc_camera* cam,
                                       const orc_render_settings* s, uint32_t sample_number, uint32_t rn_camera,
                                       uint32_t rn_path, float* fb, double* fb64, int32_t* gb_prims, int32_t nthreads)
{
    orc_render_settings plain = *s;
    if (plain.direct_light_sampling_strategy != 4)
        plain.direct_light_sampling_strategy = 3;
    orc_frame* f = orc_frame_create(sc, tree, cam, &plain, NULL);
    orc_frame_camera_pass(f, sample_number, rn_camera, nthreads);
    if (gb_prims)
        orc_frame_gbuffer_prims(f, gb_prims);
    int32_t rejected = orc_frame_path_trace(f, sample_number, rn_path, fb, fb64, nthreads);
    orc_frame_free(f);
    return rejected;
}

/* ------------------------------------------------------------------ */
/* Directional-albedo LUT baker (src/Device/kernels/Baking/ kernels,   */
/* src/Renderer/Baker/GPUBakerKernel.cpp:52-100, GPUBakerConstants.h)    */
/* ------------------------------------------------------------------ */
/* Tables: 0 GGX conductor E(cos, roughness); 1 glossy dielectric; 2 GGX
 * glass entering; 3 GGX glass exiting (relative IOR inverted); 4 GGX thin
 * glass; 5 GGX conductor weighted by the dielectric Fresnel of the
 * sampled direction. One texel integrates `launches` reference launches of
 * `ipk` samples each; launch i (1-based) seeds xorshift32 with
 * wang_hash(texel + 1) * i. Launch i is taken by lane (i - 1) % lanes of
 * the texel, each lane summing its launches in increasing order, and the
 * lanes' sums combine pairwise ((l0 + l1) + (l2 + l3)) + ...; one lane is
 * the reference's own order. */
#define BAKE_ELEMENTS_PER_LAUNCH 100000000.0f /* GPUBakerConstants.h:38 */

void orc_bake_plan(int32_t table, int32_t nx, int32_t ny, int32_t nz, int32_t count, int32_t* ipk, int32_t* launches,
                   int32_t* normaliser)
{
    /* GPUBakerKernel.cpp:77-78: the launch count from the table's texels */
    float texels = (float)(nx * ny * nz);
    int32_t k = (int32_t)floorf(hmax(1.0f, BAKE_ELEMENTS_PER_LAUNCH / texels));
    int32_t n = (int32_t)ceilf((float)count / (float)k);
    /* the kernels' own nb_samples; the Fresnel kernel derives its
     * iterations per launch from the 2D size (GGXFresnelDirectionalAlbedo.h:56) */
    int32_t kk = table == 5 ? (int32_t)floorf(hmax(1.0f, BAKE_ELEMENTS_PER_LAUNCH / (float)(nx * ny))) : k;
    *ipk = k;
    *launches = n;
    *normaliser = (int32_t)ceilf((float)count / (float)kk) * kk;
}

/* torrance_sparrow_GGX_eval_reflect<0> (Microfacet.h:103-178), isotropic,
 * the peak sampled, with the masking-shadowing flavour of the bake */
static float bake_reflect(float r, float F, v3 lv, v3 ll, v3 lh, float* pdf, int correlated)
{
    *pdf = 0.0f;
    if (smooth(r))
    {
        if (vdot(reflect(lv, V(0, 0, 1)), ll) < 0.999999f)
            return 0.0f;
        *pdf = 1.0e9f;
        return 1.0e9f * F / fabsf(ll.z);
    }
    if (ll.z < 0.0f)
        return 0.0f;
    float ax, ay;
    alphas(r, 0.0f, &ax, &ay);
    float D = ggx_D(ax, ay, lh);
    float lam_v = lambda_(ax, ay, lv);
    float G1V = 1.0f / (1.0f + lam_v);
    float Dv = vndf(D, G1V, lv, lh);
    float NoV = hmax(1.0e-3f, fabsf(lv.z)), NoL = hmax(1.0e-3f, fabsf(ll.z));
    *pdf = Dv / (4.0f * vdot(lv, lh));
    if (*pdf == 0.0f)
        return 0.0f;
    float lam_l = lambda_(ax, ay, ll);
    float G2 = correlated ? 1.0f / (1.0f + lam_v + lam_l) : G1V * (1.0f / (1.0f + lam_l));
    return F * D * G2 / (4.0f * NoL * NoV);
}

/* GGX_anisotropic_sample_microfacet (Microfacet.h:460-476) */
static v3 bake_microfacet(v3 lv, float r, uint32_t* st)
{
    float ax, ay;
    alphas(r, 0.0f, &ax, &ay);
    if (ax <= 1.0e-4f && ay <= 1.0e-4f)
        return V(0, 0, 1);
    return vndf_sample(lv, ax, ay, st);
}

/* GGX_glass_E_sample / thin_glass_sample (GGXGlassDirectionalAlbedo.h:123-160,
 * GGXThinGlassDirectionalAlbedo.h:32-86) */
static v3 bake_glass_sample(float eta, float r, v3 lv, int thin, uint32_t* st)
{
    if (fabsf(eta - 1.0f) < 1.0e-5f)
        eta = 1.0f + 1.0e-5f;
    v3 m = bake_microfacet(lv, r, st);
    float F = fresnel_d(vdot(lv, m), eta);
    if (thin && r < 0.1f)
        F += (1.0f - F) * (1.0f - F) * F / (1.0f - F * F);
    if (orc_rand01(st) < F)
        return reflect(lv, m);
    if (vdot(m, lv) < 0.0f)
        m = vneg(m);
    if (thin)
    {
        v3 d = reflect(lv, m);
        d.z *= -1.0f;
        return d;
    }
    return refract_dir(lv, m, eta);
}

/* GGX_glass_E_eval / thin_glass_eval (GGXGlassDirectionalAlbedo.h:36-118,
 * GGXThinGlassDirectionalAlbedo.h:88-189) */
static float bake_glass_eval(float eta, float r, v3 lv, v3 ll, float* pdf, int thin, int correlated)
{
    *pdf = 0.0f;
    float NoV = lv.z, NoL = ll.z;
    if (fabsf(NoL) < 1.0e-8f)
        return 0.0f;
    int reflecting = NoL * NoV > 0.0f;
    if (fabsf(eta - 1.0f) < 1.0e-5f)
        eta = 1.0f + 1.0e-5f;
    v3 h;
    if (reflecting)
        h = vadd(ll, lv);
    else if (thin)
        h = vadd(V(ll.x, ll.y, -ll.z), lv);
    else
        h = vadd(vscale(ll, eta), lv);
    h = vnorm(h);
    if (h.z < 0.0f)
        h = vneg(h);
    float HoL = vdot(ll, h), HoV = vdot(lv, h);
    if (HoL * NoL < 0.0f || HoV * NoV < 0.0f)
        return 0.0f;
    float F = fresnel_d(HoV, eta);
    if (thin && r < 0.1f)
        F += (1.0f - F) * (1.0f - F) * F / (1.0f - F * F);
    if (reflecting)
    {
        float v = bake_reflect(r, F, lv, ll, h, pdf, correlated);
        *pdf *= F;
        return v;
    }
    float dp = HoL + HoV / eta;
    float dp2 = dp * dp;
    float denom = dp2 * NoL * NoV;
    float ax, ay;
    alphas(r, 0.0f, &ax, &ay);
    float D = ggx_D(ax, ay, h);
    float G1V = 1.0f / (1.0f + lambda_(ax, ay, lv));
    float G1L = 1.0f / (1.0f + lambda_(ax, ay, ll));
    float G2 = G1V * G1L;
    float dwm_dwi = fabsf(HoL) / dp2;
    float D_pdf = G1V / fabsf(NoV) * D * fabsf(HoV);
    *pdf = dwm_dwi * D_pdf;
    *pdf *= 1.0f - F;
    return D * (1.0f - F) * G2 * fabsf(HoL * HoV / denom);
}

/* one sample of a table's integrand; 0 with *skip set for the samples the
 * reference `continue`s past */
static float bake_sample(int32_t table, float r, float eta, v3 lv, int correlated, uint32_t* st, int* skip)
{
    *skip = 0;
    float pdf;
    if (table == 0 || table == 5)
    {
        /* GGXConductorDirectionalAlbedo.h:60-78, GGXFresnelDirectionalAlbedo.h:60-77 */
        v3 ll = ggx_sample(r, 0.0f, lv, st);
        if (ll.z < 0.0f)
        {
            *skip = 1;
            return 0.0f;
        }
        float F = table == 5 ? fresnel_d(ll.z, eta) : 1.0f;
        float v = bake_reflect(r, F, lv, ll, vnorm(vadd(lv, ll)), &pdf, correlated);
        v /= pdf;
        v *= ll.z;
        return v;
    }
    if (table == 1)
    {
        /* GlossyDielectricDirectionalAlbedo.h:64-119 */
        float lobe = orc_rand01(st);
        v3 ll;
        if (lobe < 0.5f)
        {
            ll = ggx_sample(r, 0.0f, lv, st);
            if (ll.z < 0.0f)
            {
                *skip = 1;
                return 0.0f;
            }
        }
        else
        {
            /* cosine_weighted_sample_z_up_frame (Sampling.h:210-220) */
            float r1 = orc_rand01(st), r2 = orc_rand01(st);
            float phi = M_TWO_PI_F * r1, ct = sqrtf(r2), stt = sqrtf(1.0f - ct * ct);
            float s, c;
            orc_sincos(phi, &s, &c);
            ll = vnorm(V(c * stt, s * stt, ct));
        }
        v3 h = vnorm(vadd(lv, ll));
        float F = fresnel_d(vdot(h, ll), eta);
        float pdf_s;
        float spec = bake_reflect(r, F, lv, ll, h, &pdf_s, correlated);
        float total = pdf_s * 0.5f;
        float thr = 1.0f;
        thr *= 1.0f - fresnel_d(ll.z, eta);
        thr *= 1.0f - fresnel_d(lv.z, eta);
        /* lambertian_brdf_eval (Lambertian.h:14-23) of a white base colour */
        float pdf_d = 0.0f, diff = 0.0f;
        if (ll.z > 0.0f)
        {
            pdf_d = ll.z * M_INV_PI_F;
            diff = M_INV_PI_F;
        }
        total += pdf_d * 0.5f;
        diff *= thr;
        float v = spec + diff;
        v *= ll.z;
        v /= total;
        return v;
    }
    /* glass, its inverse and thin glass: GGXGlassDirectionalAlbedo.h:191-203,
     * GGXThinGlassDirectionalAlbedo.h:231-244 */
    int thin = table == 4;
    float rr = thin ? thin_rough(1, r, eta) : r;
    v3 ll = bake_glass_sample(eta, rr, lv, thin, st);
    float v = bake_glass_eval(eta, rr, lv, ll, &pdf, thin, correlated);
    if (pdf == 0.0f)
    {
        *skip = 1;
        return 0.0f;
    }
    v /= pdf;
    v *= fabsf(ll.z);
    return v;
}

/* the texel parameters (GGXConductorDirectionalAlbedo.h:44-52,
 * GlossyDielectricDirectionalAlbedo.h:46-63): cos_theta_o, roughness and
 * the relative IOR from F0 = (z / (nz - 1))^4 */
static void bake_texel_params(int32_t table, int32_t nx, int32_t ny, int32_t nz, int32_t x, int32_t y, int32_t z,
                              float* r_out, float* eta_out, v3* lv_out)
{
    float cos_o = 1.0f / ((float)nx - 1.0f) * (float)x;
    cos_o = hmax(1.0e-3f, cos_o);
    if (table != 0 && table != 4)
        cos_o = cos_o * cos_o * sqrtf(cos_o); /* powf(cos_o, 2.5f) */
    float sin_o = orc_sinf(orc_acosf(cos_o));
    float r = 1.0f / ((float)ny - 1.0f) * (float)y;
    r = hmax(r, 1.0e-4f);
    float eta = 1.0f;
    if (table != 0)
    {
        float F0 = 1.0f / ((float)nz - 1.0f) * (float)z;
        F0 *= F0;
        F0 *= F0;
        float sq = sqrtf(hclamp(0.0f, 0.99f, F0));
        eta = (1.0f + sq) / (1.0f - sq);
        if (table == 3)
            eta = 1.0f / eta;
    }
    *r_out = r;
    *eta_out = eta;
    *lv_out = vnorm(V(sin_o, 0.0f, cos_o));
}

float orc_bake_texel(int32_t table, int32_t nx, int32_t ny, int32_t nz, int32_t count, int32_t correlated,
                     int32_t lanes, int32_t texel)
{
    int32_t ipk, launches, norm;
    orc_bake_plan(table, nx, ny, nz, count, &ipk, &launches, &norm);
    int32_t x = texel % nx, y = (texel / nx) % ny, z = texel / (nx * ny);
    float r, eta;
    v3 lv;
    bake_texel_params(table, nx, ny, nz, x, y, z, &r, &eta, &lv);
    float part[64];
    for (int32_t j = 0; j < lanes; j++)
    {
        float acc = 0.0f;
        for (int32_t i = j + 1; i <= launches; i += lanes)
        {
            uint32_t st = orc_wang_hash((uint32_t)texel + 1u) * (uint32_t)i;
            for (int32_t s = 0; s < ipk; s++)
            {
                int skip;
                float v = bake_sample(table, r, eta, lv, correlated, &st, &skip);
                /* a NaN sample (thin-glass transmission whose generalized
                 * half-vector term HoL + HoV / eta cancels to exactly 0 in
                 * binary32: inf / inf) is skipped like a zero-pdf one */
                if (!skip && v == v)
                    acc += v / (float)norm;
            }
        }
        part[j] = acc;
    }
    for (int32_t off = 1; off < lanes; off <<= 1)
        for (int32_t j = 0; j < lanes; j += 2 * off)
            part[j] = part[j] + part[j + off];
    return part[0];
}

void orc_bake_texels(int32_t table, int32_t nx, int32_t ny, int32_t nz, int32_t count, int32_t correlated,
                     int32_t lanes, const int32_t* texels, int32_t n, float* out, int32_t nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int32_t i = 0; i < n; i++)
        out[i] = orc_bake_texel(table, nx, ny, nz, count, correlated, lanes, texels[i]);
}
