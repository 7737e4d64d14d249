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
    int do_metallic_ec;
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

static omat unpack_material(const orc_cpu_material* m)
{
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
    const float* lut;
    const orc_octree* tree;
    const orc_camera* cam;
    orc_render_settings s;
    uint32_t sample_number;
    float consistent_tau;
} rctx;

enum { NO_INFO = 0, FROM_COAT = 2, FROM_FIRST_METAL = 4, FROM_SECOND_METAL = 8, FROM_SPECULAR = 16,
       FROM_DIFFUSE = 128, FROM_DIFFUSE_TRANSMISSION = 256 };
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
    int first = info == FROM_FIRST_METAL && smooth(m->roughness) && matching;
    int second = info == FROM_SECOND_METAL && smooth(m->second_roughness) && matching;
    return (first || second) ? 1 : 2;
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
typedef struct { v3 view, sn, gn, L; int info, reg; float path_r; } bctx;

static rgb metallic_eval(const rctx* R, const omat* m, const bctx* c, float r, v3 lv, v3 ll, v3 lh, float* pdf)
{
    float rr = regularize(R, c->reg, r, c->path_r);
    int ds = delta_sampled(m, rr, m->anisotropy, c->info);
    if (ds == 2)
    {
        *pdf = 0.0f;
        return C1(0.0f);
    }
    float HoL = hclamp(1.0e-8f, 1.0f, vdot(lh, ll));
    rgb F = f82(m->base_color, m->F82, m->F90, m->F90_exp, HoL);
    rgb ms = ec_conductor(R, F, rr, m->do_metallic_ec, lv);
    rgb ss = ggx_reflect_ss(rr, m->anisotropy, F, lv, ll, lh, pdf, ds);
    return cmul(ss, ms);
}

typedef struct { float coat, sheen, m1, m2, spec, diff, glass, dt; } lobes;

static lobes weights(const omat* m)
{
    lobes w;
    const float outside = 1.0f;
    w.coat = m->coat * outside;
    w.sheen = m->sheen * outside;
    w.m1 = m->metallic * outside;
    w.m2 = m->metallic * outside;
    w.m1 = hlerp(w.m1, 0.0f, m->second_roughness_weight);
    w.m2 = hlerp(0.0f, w.m2, m->second_roughness_weight);
    w.glass = (1.0f - m->metallic) * (1.0f - m->diffuse_transmission) * m->specular_transmission;
    w.dt = (1.0f - m->metallic) * m->diffuse_transmission;
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

static rgb principled_eval(const rctx* R, const omat* m, const bctx* c, float* pdf)
{
    *pdf = 0.0f;
    int refracting = vdot(c->sn, c->L) < 0.0f;
    v3 T, B, TR, BR;
    build_onb(c->sn, &T, &B);
    v3 lv = to_local(T, B, c->sn, c->view), ll = to_local(T, B, c->sn, c->L);
    build_rotated_onb(c->sn, &TR, &BR, m->anisotropy_rotation * M_PI_F);
    v3 lvr = to_local(TR, BR, c->sn, c->view), llr = to_local(TR, BR, c->sn, c->L);
    v3 lhr = vnorm(vadd(lvr, llr));
    lobes w = weights(m), p = probas(w);
    rgb layers = C1(1.0f), out = C1(0.0f);
    float wm[2] = {w.m1 * (float)!refracting, w.m2 * (float)!refracting};
    float pm[2] = {p.m1, p.m2}, rm[2] = {m->roughness, m->second_roughness};
    for (int k = 0; k < 2; k++)
    {
        rgb contrib = C1(0.0f);
        if (wm[k] > 0.0f && lvr.z > 0.0f && llr.z > 0.0f)
        {
            float mp = 0.0f;
            contrib = metallic_eval(R, m, c, rm[k], lvr, llr, lhr, &mp);
            contrib = cscale(contrib, wm[k]);
            contrib = cmul(contrib, layers);
            *pdf += mp * pm[k];
        }
        out = cadd(out, contrib);
    }
    rgb glossy = cadd(C1(0.0f), C1(0.0f));
    float wd = w.diff * (float)!refracting;
    if (wd > 0.0f && lv.z > 0.0f && ll.z > 0.0f)
    {
        /* lambertian_brdf_eval (Lambertian.h:14-23) */
        float dp = ll.z * M_INV_PI_F;
        rgb contrib = cscale(m->base_color, M_INV_PI_F);
        contrib = cscale(contrib, wd);
        contrib = cmul(contrib, layers);
        *pdf += dp * p.diff;
        glossy = cadd(glossy, contrib);
    }
    out = cadd(out, cdivs(glossy, 1.0f));
    return cdivs(out, 1.0f);
}

static rgb principled_sample(const rctx* R, const omat* m, bctx* c, v3* dir, float* pdf, uint32_t* st)
{
    *pdf = 0.0f;
    lobes w = weights(m), p = probas(w);
    float cdf0 = p.coat, cdf1 = cdf0 + p.sheen, cdf2 = cdf1 + p.m1, cdf3 = cdf2 + p.m2;
    float cdf4 = cdf3 + p.spec, cdf5 = cdf4 + p.diff, cdf6 = cdf5 + p.dt;
    float r1 = orc_rand01(st);
    int s_dt = r1 > cdf5 && r1 < cdf6, s_glass = r1 > cdf6;
    v3 TR, BR;
    build_rotated_onb(c->sn, &TR, &BR, m->anisotropy_rotation * M_PI_F);
    v3 lvr = to_local(TR, BR, c->sn, c->view);
    if (r1 < cdf0) return C1(0.0f);
    else if (r1 < cdf1) return C1(0.0f);
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
    else if (r1 < cdf4) return C1(0.0f);
    else if (r1 < cdf5)
    {
        c->info = FROM_DIFFUSE;
        *dir = cosine_around_normal(c->sn, st);
    }
    else return C1(0.0f);
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

static int trace_main(const rctx* R, v3 o, v3 d, int32_t last, hitinfo* out)
{
    orc_hit h = trace1(R, o, d, last, INFINITY);
    if (h.prim < 0)
        return 0;
    v3 a, b, c;
    tri(R, h.prim, &a, &b, &c);
    out->p = vadd(o, vscale(d, h.t));
    out->gn = vnorm(vnorm(vcross(vsub(b, a), vsub(c, a))));
    out->sn = out->gn;
    out->prim = h.prim;
    fix_backfacing(vneg(d), &out->gn, &out->sn);
    return 1;
}

/* ------------------------------------------------------------------ */
/* NEE with MIS (Lights.h:123-226, LightUtils.h:142-277,               */
/* PDFConversion.h:19-38, Sampling.h:103-136)                           */
/* ------------------------------------------------------------------ */
static rgb sample_one_light_mis(const rctx* R, const omat* m, const hitinfo* hi, v3 view, float path_r, uint32_t* st)
{
    rgb light_mis = C1(0.0f);
    {
        int32_t e = orc_random_index(st, R->ne);
        int32_t lt = R->emissive[e];
        v3 A, B, Cc;
        tri(R, lt, &A, &B, &Cc);
        float r1 = orc_rand01(st), r2 = orc_rand01(st);
        if (r2 > r1) { r1 *= 0.5f; r2 -= r1; }
        else { r2 *= 0.5f; r1 -= r2; }
        v3 AB = vsub(B, A), AC = vsub(Cc, A);
        v3 n = vcross(AB, AC);
        float len = vlen(n);
        v3 point = vadd(vadd(A, vscale(AB, r1)), vscale(AC, r2));
        v3 ln = vdivs(n, len);
        float area = 0.5f * len;
        float apdf = 1.0f / area;
        apdf /= (float)R->ne;
        rgb emission = R->mats[R->mat_idx[lt]].emission;
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
                bctx c = {view, hi->sn, hi->gn, dn, NO_INFO, REG_MIS, path_r};
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
        bctx c = {view, hi->sn, hi->gn, V(0, 0, 0), NO_INFO, REG_MIS, path_r};
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
                    float pa = 1.0f / R->areas[h.prim];
                    pa /= (float)R->ne;
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
    for (int bounce = 0; bounce < R->s.nb_bounces + 1; bounce++)
    {
        if (missed)
            break;
        if (bounce > 0)
        {
            v3 o = hi.p;
            int32_t last = hi.prim;
            found = trace_main(R, o, dir, last, &hi);
            if (found)
                mat = R->mat_idx[hi.prim];
        }
        if (!found)
        {
            missed = 1; /* no envmap in these scenes: adds 0 */
            continue;
        }
        const omat* m = &R->mats[mat];
        v3 view = vneg(dir);
        if (bounce > 0 || R->s.enable_direct)
        {
            /* estimate_direct_lighting (Lights.h:364-398) */
            rgb direct = C1(0.0f);
            if (R->ne > 0)
                direct = cadd(C1(0.0f), sample_one_light_mis(R, m, &hi, view, path_r, &st));
            rgb total = C1(0.0f);
            if (bounce == 0)
                total = cadd(total, m->emission);
            total = cadd(total, cmul(cadd(cdivs(direct, 1.0f), C1(0.0f)), thr));
            color = cadd(color, total);
        }
        /* path_tracing_compute_next_indirect_bounce (PathTracing.h:26-101) */
        bctx c = {view, hi.sn, hi.gn, V(0, 0, 0), NO_INFO, NO_REG, path_r};
        v3 nd;
        float bpdf;
        rgb f = principled_sample(R, m, &c, &nd, &bpdf, &st);
        if (c.info == FROM_DIFFUSE || c.info == FROM_DIFFUSE_TRANSMISSION)
            path_r = 1.0f;
        else if (c.info == FROM_FIRST_METAL || c.info == FROM_SPECULAR)
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
        thr = cmul(thr, C1(1.0f));
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

int32_t orc_render_sample(const orc_render_scene* sc, const orc_octree* tree, const orc_camera* cam,
                          const orc_render_settings* s, uint32_t sample_number, uint32_t rn_camera,
                          uint32_t rn_path, float* fb, double* fb64, int32_t* gb_prims, int32_t nthreads)
{
    omat* mats = (omat*)malloc(sizeof(omat) * (size_t)(sc->material_count > 0 ? sc->material_count : 1));
    for (int32_t i = 0; i < sc->material_count; i++)
        mats[i] = unpack_material(sc->materials + i);
    float* areas = (float*)malloc(sizeof(float) * (size_t)(sc->triangle_count > 0 ? sc->triangle_count : 1));
    for (int32_t t = 0; t < sc->triangle_count; t++)
    {
        /* ThreadFunctions.cpp:221-242 */
        const float* a = sc->vertices + 3 * (int64_t)sc->indices[3 * (int64_t)t];
        const float* b = sc->vertices + 3 * (int64_t)sc->indices[3 * (int64_t)t + 1];
        const float* c = sc->vertices + 3 * (int64_t)sc->indices[3 * (int64_t)t + 2];
        v3 n = vcross(V(b[0] - a[0], b[1] - a[1], b[2] - a[2]), V(c[0] - a[0], c[1] - a[1], c[2] - a[2]));
        areas[t] = vlen(n) * 0.5f;
    }
    rctx R = {sc->vertices, sc->indices, sc->material_indices, mats, sc->emissive, sc->emissive_count, areas,
              sc->lut_conductor, tree, cam, *s, sample_number, 0.0f};
    R.consistent_tau = orc_consistent_tau(s->microfacet_regularization_tau_0, sample_number);
    int32_t rejected = 0;
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic) reduction(+ : rejected)
    for (int y = 0; y < s->height; y++)
    {
        for (int x = 0; x < s->width; x++)
        {
            size_t pix = (size_t)x + (size_t)y * (size_t)s->width;
            gbuf g;
            camera_rays_pixel(&R, x, y, rn_camera, &g);
            if (gb_prims)
                gb_prims[pix] = g.prim;
            rgb c = C1(0.0f);
            if (g.prim != -1)
                c = megakernel_pixel(&R, x, y, rn_path, &g);
            /* SanityCheck.h:78-106: drop NaN/inf/negative samples */
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
            float* f = fb + 3 * pix;
            if (sample_number == 0)
            {
                f[0] = c.r; f[1] = c.g; f[2] = c.b;
            }
            else
            {
                f[0] += c.r; f[1] += c.g; f[2] += c.b;
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
    }
    free(mats);
    free(areas);
    return rejected;
}

/* ------------------------------------------------------------------ */
/* BSDF entry points for the oracle's own property tests.               */
/* ------------------------------------------------------------------ */
static rctx bsdf_ctx(const float* lut, float consistent_tau, int32_t regularize)
{
    rctx R;
    memset(&R, 0, sizeof(R));
    R.lut = lut;
    R.consistent_tau = consistent_tau;
    R.s.do_microfacet_regularization = regularize;
    return R;
}

void orc_bsdf_eval(const orc_cpu_material* cm, const float* lut, float consistent_tau, int32_t reg_mode,
                   float path_roughness, const float* view, const float* sn, const float* gn, const float* to_light,
                   float* out_rgb, float* out_pdf)
{
    rctx R = bsdf_ctx(lut, consistent_tau, 1);
    omat m = unpack_material(cm);
    bctx c = {V(view[0], view[1], view[2]), V(sn[0], sn[1], sn[2]), V(gn[0], gn[1], gn[2]),
              V(to_light[0], to_light[1], to_light[2]), NO_INFO, reg_mode, path_roughness};
    rgb f = principled_eval(&R, &m, &c, out_pdf);
    out_rgb[0] = f.r; out_rgb[1] = f.g; out_rgb[2] = f.b;
}

void orc_bsdf_sample(const orc_cpu_material* cm, const float* lut, float consistent_tau, int32_t reg_mode,
                     float path_roughness, const float* view, const float* sn, const float* gn, uint32_t* rng_state,
                     float* out_dir, float* out_rgb, float* out_pdf)
{
    rctx R = bsdf_ctx(lut, consistent_tau, 1);
    omat m = unpack_material(cm);
    bctx c = {V(view[0], view[1], view[2]), V(sn[0], sn[1], sn[2]), V(gn[0], gn[1], gn[2]), V(0, 0, 0), NO_INFO,
              reg_mode, path_roughness};
    v3 d = V(0, 0, 0);
    rgb f = principled_sample(&R, &m, &c, &d, out_pdf, rng_state);
    out_dir[0] = d.x; out_dir[1] = d.y; out_dir[2] = d.z;
    out_rgb[0] = f.r; out_rgb[1] = f.g; out_rgb[2] = f.b;
}
