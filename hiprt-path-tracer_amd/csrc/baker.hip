// baker.hip -- the energy-compensation LUT baker (GPUBaker, src/Renderer/Baker/
// GPUBaker.cpp:37-93, kernels src/Device/kernels/Baking/) on the MI355X.
//
// A texel's value is a Monte Carlo directional albedo over `launches` x
// `ipk` samples: the reference's host loop runs `launches` kernel launches
// of `ipk` samples per texel (GPUBakerKernel.cpp:77-92, 1e8 texel-samples a
// launch), launch i seeding the texel's xorshift32 with wang_hash(texel + 1)
// * i. Here the launches of a texel are dealt to `lanes` consecutive lanes of
// a wave (lane j takes launches j + 1, j + 1 + lanes, ...), each summing its
// samples in the reference's order, and the lanes' partial sums combine
// pairwise through cross-lane shuffles. One lane (the default) is the
// reference's order, which matters: a sequential binary32 sum of 262,144
// samples of ~4e-6 drifts by up to ~0.2 % (the conductor's smooth rows come
// out at 1.0009), comparable to the 8-bit mantissa of the shipped RGBE
// tables, so only that order reproduces them. More lanes fill the 256 CUs
// on small tables (16,384 conductor texels are 256 waves) and sum more
// accurately. The work is ALU-bound: 4 bytes leave the chip per texel.
#include <cmath>
#include <cstring>

#include "internal.h"
#include "shading.h"

namespace mipt {

namespace {

using namespace pt;

struct BakeParams {
    int32_t nx, ny, nz;
    int32_t ipk;          // samples per texel per reference launch
    int32_t launches;     // reference launches
    int32_t lane_bits;    // log2(lanes per texel)
    float normalizer;     // nb_samples dividing each sample
    uint32_t texels;
    int32_t m_begin, m_end; // this kernel launch runs each lane's launches m_begin .. m_end - 1
    int32_t first, last;    // the partial sums start at 0 / are reduced and written
};

// torrance_sparrow_GGX_eval_reflect<0> (Microfacet.h:103-178) of an isotropic
// lobe whose peak was sampled, with the masking-shadowing flavour of the bake.
template <bool CORRELATED>
__device__ __forceinline__ float bake_reflect(float r, float F, f3 lv, f3 ll, f3 lh, float& pdf)
{
    pdf = 0.0f;
    if (is_perfectly_smooth(r))
    {
        if (dot(reflect_ray(lv, mk(0.0f, 0.0f, 1.0f)), ll) < DELTA_ALIGN_THRESHOLD)
            return 0.0f;
        pdf = DELTA_HIGH_VALUE;
        return DELTA_HIGH_VALUE * F / fabsf(ll.z);
    }
    if (ll.z < 0.0f)
        return 0.0f;
    float ax, ay;
    get_alphas(r, 0.0f, ax, ay);
    const float D = ggx_d(ax, ay, lh);
    const float lam_v = smith_lambda(ax, ay, lv);
    const float G1V = 1.0f / (1.0f + lam_v);
    const float Dv = ggx_vndf(D, G1V, lv, lh);
    const float NoV = maxf_(GGX_DOT_CLAMP, fabsf(lv.z));
    const float NoL = maxf_(GGX_DOT_CLAMP, fabsf(ll.z));
    pdf = Dv / (4.0f * dot(lv, lh));
    if (pdf == 0.0f)
        return 0.0f;
    const float lam_l = smith_lambda(ax, ay, ll);
    const float G2 = CORRELATED ? 1.0f / (1.0f + lam_v + lam_l) : G1V * (1.0f / (1.0f + lam_l));
    return F * D * G2 / (4.0f * NoL * NoV);
}

// GGX_anisotropic_sample_microfacet (Microfacet.h:460-476), isotropic.
__device__ __forceinline__ f3 bake_microfacet(f3 lv, float r, Xorshift32& rng)
{
    float ax, ay;
    get_alphas(r, 0.0f, ax, ay);
    if (ax <= ROUGHNESS_CLAMP && ay <= ROUGHNESS_CLAMP)
        return mk(0.0f, 0.0f, 1.0f);
    return ggx_vndf_sample(lv, ax, ay, rng);
}

// Thin-walled Fresnel boost below roughness 0.1 (GGXThinGlassDirectionalAlbedo.h:52-53).
__device__ __forceinline__ float thin_boost(float F) { return F + (1.0f - F) * (1.0f - F) * F / (1.0f - F * F); }

// GGX_glass_E_sample / thin_glass_sample (GGXGlassDirectionalAlbedo.h:123-160,
// GGXThinGlassDirectionalAlbedo.h:32-86).
template <bool THIN>
__device__ __forceinline__ f3 bake_glass_sample(float eta, float r, f3 lv, Xorshift32& rng)
{
    if (fabsf(eta - 1.0f) < 1.0e-5f)
        eta = 1.0f + 1.0e-5f;
    f3 m = bake_microfacet(lv, r, rng);
    float F = fresnel_dielectric(dot(lv, m), eta);
    if (THIN && r < 0.1f)
        F = thin_boost(F);
    if (rng() < F)
        return reflect_ray(lv, m);
    if (dot(m, lv) < 0.0f)
        m = -m;
    if (THIN)
    {
        f3 d = reflect_ray(lv, m);
        d.z *= -1.0f;
        return d;
    }
    return refract_ray(lv, m, eta);
}

// GGX_glass_E_eval / thin_glass_eval (GGXGlassDirectionalAlbedo.h:36-118,
// GGXThinGlassDirectionalAlbedo.h:88-189).
template <bool THIN, bool CORRELATED>
__device__ __forceinline__ float bake_glass_eval(float eta, float r, f3 lv, f3 ll, float& pdf)
{
    pdf = 0.0f;
    const float NoV = lv.z, NoL = ll.z;
    if (fabsf(NoL) < 1.0e-8f)
        return 0.0f;
    const bool reflecting = NoL * NoV > 0.0f;
    if (fabsf(eta - 1.0f) < 1.0e-5f)
        eta = 1.0f + 1.0e-5f;
    f3 h;
    if (reflecting)
        h = ll + lv;
    else if (THIN)
        h = mk(ll.x, ll.y, -ll.z) + lv;
    else
        h = eta * ll + lv;
    h = normalize(h);
    if (h.z < 0.0f)
        h = -h;
    const float HoL = dot(ll, h), HoV = dot(lv, h);
    if (HoL * NoL < 0.0f || HoV * NoV < 0.0f)
        return 0.0f;
    float F = fresnel_dielectric(HoV, eta);
    if (THIN && r < 0.1f)
        F = thin_boost(F);
    if (reflecting)
    {
        const float v = bake_reflect<CORRELATED>(r, F, lv, ll, h, pdf);
        pdf *= F;
        return v;
    }
    const float dp = HoL + HoV / eta;
    const float dp2 = dp * dp;
    const float denom = dp2 * NoL * NoV;
    float ax, ay;
    get_alphas(r, 0.0f, ax, ay);
    const float D = ggx_d(ax, ay, h);
    const float G1V = 1.0f / (1.0f + smith_lambda(ax, ay, lv));
    const float G1L = 1.0f / (1.0f + smith_lambda(ax, ay, ll));
    const float G2 = G1V * G1L;
    const float dwm_dwi = fabsf(HoL) / dp2;
    const float D_pdf = G1V / fabsf(NoV) * D * fabsf(HoV);
    pdf = dwm_dwi * D_pdf;
    pdf *= 1.0f - F;
    return D * (1.0f - F) * G2 * fabsf(HoL * HoV / denom);
}

// One sample of a table's integrand; `skip` for the samples the reference
// `continue`s past.
template <int TABLE, bool CORRELATED>
__device__ __forceinline__ float bake_sample(float r, float eta, f3 lv, Xorshift32& rng, bool& skip)
{
    skip = false;
    float pdf;
    if constexpr (TABLE == MIPT_LUT_GGX_CONDUCTOR || TABLE == MIPT_LUT_GGX_FRESNEL)
    {
        // GGXConductorDirectionalAlbedo.h:60-78, GGXFresnelDirectionalAlbedo.h:60-77
        const f3 ll = ggx_sample_reflection(r, 0.0f, lv, rng);
        if (ll.z < 0.0f)
        {
            skip = true;
            return 0.0f;
        }
        const float F = TABLE == MIPT_LUT_GGX_FRESNEL ? fresnel_dielectric(ll.z, eta) : 1.0f;
        float v = bake_reflect<CORRELATED>(r, F, lv, ll, normalize(lv + ll), pdf);
        v /= pdf;
        v *= ll.z;
        return v;
    }
    else if constexpr (TABLE == MIPT_LUT_GLOSSY_DIELECTRIC)
    {
        // GlossyDielectricDirectionalAlbedo.h:64-119
        const float lobe = rng();
        f3 ll;
        if (lobe < 0.5f)
        {
            ll = ggx_sample_reflection(r, 0.0f, lv, rng);
            if (ll.z < 0.0f)
            {
                skip = true;
                return 0.0f;
            }
        }
        else
            ll = cosine_sample_z_up(rng);
        const f3 h = normalize(lv + ll);
        const float F = fresnel_dielectric(dot(h, ll), eta);
        float pdf_s;
        const float spec = bake_reflect<CORRELATED>(r, F, lv, ll, h, pdf_s);
        float total = pdf_s * 0.5f;
        float thr = 1.0f;
        thr *= 1.0f - fresnel_dielectric(ll.z, eta);
        thr *= 1.0f - fresnel_dielectric(lv.z, eta);
        // lambertian_brdf_eval (Lambertian.h:14-23) of a white base colour
        float pdf_d = 0.0f, diff = 0.0f;
        if (ll.z > 0.0f)
        {
            pdf_d = ll.z * INV_PI_F;
            diff = INV_PI_F;
        }
        total += pdf_d * 0.5f;
        diff *= thr;
        float v = spec + diff;
        v *= ll.z;
        v /= total;
        return v;
    }
    else
    {
        // GGXGlassDirectionalAlbedo.h:191-203, GGXThinGlassDirectionalAlbedo.h:231-244
        constexpr bool THIN = TABLE == MIPT_LUT_GGX_THIN_GLASS;
        const float rr = THIN ? thin_walled_roughness(true, r, eta) : r;
        const f3 ll = bake_glass_sample<THIN>(eta, rr, lv, rng);
        float v = bake_glass_eval<THIN, CORRELATED>(eta, rr, lv, ll, pdf);
        if (pdf == 0.0f)
        {
            skip = true;
            return 0.0f;
        }
        v /= pdf;
        v *= fabsf(ll.z);
        return v;
    }
}

// Threads (texel, lane), `lanes` consecutive lanes of a wave per texel; a
// texel's lanes are all in range or all out. Each kernel launch runs every
// lane's launches m_begin .. m_end - 1 (the m-th is reference launch
// lane + 1 + m * lanes), carrying the partial sum in `partial` between
// kernel launches; the last one reduces and writes the texel.
template <int TABLE, bool CORRELATED>
__global__ __launch_bounds__(256) void k_bake(BakeParams p, float* __restrict__ partial, float* __restrict__ out)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t texel = t >> p.lane_bits;
    if (texel >= p.texels)
        return;
    const int32_t lanes = 1 << p.lane_bits;
    const int32_t lane = (int32_t)(t & (uint32_t)(lanes - 1));
    const int32_t x = (int32_t)(texel % (uint32_t)p.nx);
    const int32_t y = (int32_t)((texel / (uint32_t)p.nx) % (uint32_t)p.ny);
    const int32_t z = (int32_t)(texel / ((uint32_t)p.nx * (uint32_t)p.ny));

    // texel parameters (GGXConductorDirectionalAlbedo.h:44-52,
    // GlossyDielectricDirectionalAlbedo.h:46-63)
    float cos_o = 1.0f / ((float)p.nx - 1.0f) * (float)x;
    cos_o = maxf_(GGX_DOT_CLAMP, cos_o);
    if (TABLE != MIPT_LUT_GGX_CONDUCTOR && TABLE != MIPT_LUT_GGX_THIN_GLASS)
        cos_o = cos_o * cos_o * sqrtf(cos_o); // powf(cos_o, 2.5f)
    const float sin_o = sin_fixed(acos_fixed(cos_o));
    float r = 1.0f / ((float)p.ny - 1.0f) * (float)y;
    r = maxf_(r, 1.0e-4f);
    float eta = 1.0f;
    if (TABLE != MIPT_LUT_GGX_CONDUCTOR)
    {
        float F0 = 1.0f / ((float)p.nz - 1.0f) * (float)z;
        F0 *= F0;
        F0 *= F0;
        const float sq = sqrtf(clampf_(0.0f, 0.99f, F0));
        eta = (1.0f + sq) / (1.0f - sq);
        if (TABLE == MIPT_LUT_GGX_GLASS_INVERSE)
            eta = 1.0f / eta;
    }
    const f3 lv = normalize(mk(sin_o, 0.0f, cos_o));

    const uint32_t hash = wang_hash(texel + 1u);
    float acc = p.first ? 0.0f : partial[t];
    for (int32_t m = p.m_begin; m < p.m_end; m++)
    {
        const int32_t i = lane + 1 + m * lanes;
        if (i > p.launches)
            break;
        Xorshift32 rng{hash * (uint32_t)i};
        for (int32_t s = 0; s < p.ipk; s++)
        {
            bool skip;
            const float v = bake_sample<TABLE, CORRELATED>(r, eta, lv, rng, skip);
            // a NaN sample (thin-glass transmission whose HoL + HoV / eta
            // cancels to exactly 0 in binary32: inf / inf) is skipped like a
            // zero-pdf one
            if (!skip && v == v)
                acc += v / p.normalizer;
        }
    }
    if (!p.last)
    {
        partial[t] = acc;
        return;
    }
    for (int32_t off = 1; off < lanes; off <<= 1)
        acc += __shfl_xor(acc, off, 64);
    if (lane == 0)
        out[(size_t)x + (size_t)(p.ny - 1 - y) * (size_t)p.nx + (size_t)z * (size_t)p.nx * (size_t)p.ny] = acc;
}

template <int TABLE>
void launch_bake(bool correlated, dim3 grid, const BakeParams& p, float* partial, float* out)
{
    if (correlated)
        k_bake<TABLE, true><<<grid, 256>>>(p, partial, out);
    else
        k_bake<TABLE, false><<<grid, 256>>>(p, partial, out);
}

void launch_table(int32_t table, bool correlated, dim3 grid, const BakeParams& p, float* partial, float* out)
{
    switch (table)
    {
    case MIPT_LUT_GGX_CONDUCTOR: launch_bake<MIPT_LUT_GGX_CONDUCTOR>(correlated, grid, p, partial, out); break;
    case MIPT_LUT_GLOSSY_DIELECTRIC: launch_bake<MIPT_LUT_GLOSSY_DIELECTRIC>(correlated, grid, p, partial, out); break;
    case MIPT_LUT_GGX_GLASS: launch_bake<MIPT_LUT_GGX_GLASS>(correlated, grid, p, partial, out); break;
    case MIPT_LUT_GGX_GLASS_INVERSE: launch_bake<MIPT_LUT_GGX_GLASS_INVERSE>(correlated, grid, p, partial, out); break;
    case MIPT_LUT_GGX_THIN_GLASS: launch_bake<MIPT_LUT_GGX_THIN_GLASS>(correlated, grid, p, partial, out); break;
    default: launch_bake<MIPT_LUT_GGX_FRESNEL>(correlated, grid, p, partial, out); break;
    }
}

// Samples one kernel launch evaluates at most (about 0.1 s of the chip).
constexpr double SAMPLES_PER_KERNEL = 4.0e9;
// Threads a table needs to fill the chip: 4 waves on each of the 1024 SIMDs.
constexpr uint32_t FILL_THREADS = 1u << 18;

} // namespace

int32_t bake_default_settings(int32_t table, mipt_bake_settings* s)
{
    std::memset(s, 0, sizeof(*s));
    s->table = table;
    s->masking_shadowing = MIPT_GGX_HEIGHT_UNCORRELATED;
    s->lanes_per_texel = 1;
    switch (table)
    {
    case MIPT_LUT_GGX_CONDUCTOR: // GGXConductorDirectionalAlbedoSettings.h, GPUBakerConstants.h:16-17
        s->size_cos_theta = 128, s->size_roughness = 128, s->size_ior = 1, s->integration_sample_count = 262144;
        return MIPT_OK;
    case MIPT_LUT_GLOSSY_DIELECTRIC: // GlossyDielectricDirectionalAlbedoSettings.h, :31-33
        s->size_cos_theta = 128, s->size_roughness = 64, s->size_ior = 128, s->integration_sample_count = 131072;
        return MIPT_OK;
    case MIPT_LUT_GGX_GLASS:
    case MIPT_LUT_GGX_GLASS_INVERSE: // GGXGlassDirectionalAlbedoSettings.h, :23-25
        s->size_cos_theta = 256, s->size_roughness = 16, s->size_ior = 128, s->integration_sample_count = 65536;
        return MIPT_OK;
    case MIPT_LUT_GGX_THIN_GLASS: // GGXThinGlassDirectionalAlbedoSettings.h, :27-29
        s->size_cos_theta = 32, s->size_roughness = 32, s->size_ior = 96, s->integration_sample_count = 65536;
        return MIPT_OK;
    case MIPT_LUT_GGX_FRESNEL: // GGXFresnelDirectionalAlbedoSettings.h, :19-21
        s->size_cos_theta = 256, s->size_roughness = 256, s->size_ior = 256, s->integration_sample_count = 65536;
        return MIPT_OK;
    default:
        return MIPT_ERR_INVALID_ARGUMENT;
    }
}

// The reference's launch plan (GPUBakerKernel.cpp:77-78) and the kernels'
// normalising count (GGXConductorDirectionalAlbedo.h:55-57; the Fresnel
// kernel derives its iterations per launch from the 2D size,
// GGXFresnelDirectionalAlbedo.h:56).
void bake_plan(const mipt_bake_settings& s, int32_t& ipk, int32_t& launches, int32_t& normalizer)
{
    const float per_launch = 100000000.0f;
    const float texels = (float)(s.size_cos_theta * s.size_roughness * s.size_ior);
    ipk = (int32_t)std::floor(std::fmax(1.0f, per_launch / texels));
    launches = (int32_t)std::ceil((float)s.integration_sample_count / (float)ipk);
    const int32_t kk = s.table == MIPT_LUT_GGX_FRESNEL
                           ? (int32_t)std::floor(std::fmax(1.0f, per_launch / (float)(s.size_cos_theta * s.size_roughness)))
                           : ipk;
    normalizer = (int32_t)std::ceil((float)s.integration_sample_count / (float)kk) * kk;
}

void bake_lut(Ctx& ctx, const mipt_bake_settings& s, float* out, mipt_bake_info* info)
{
    (void)ctx;
    if (s.table < MIPT_LUT_GGX_CONDUCTOR || s.table > MIPT_LUT_GGX_FRESNEL)
        fail(MIPT_ERR_INVALID_ARGUMENT, "bake: unknown table");
    const bool is_2d = s.table == MIPT_LUT_GGX_CONDUCTOR;
    if (s.size_cos_theta < 2 || s.size_roughness < 2 || (is_2d ? s.size_ior != 1 : s.size_ior < 2))
        fail(MIPT_ERR_INVALID_ARGUMENT,
             "bake: sizes must be >= 2 (size_ior 1 for the conductor table, >= 2 for the others)");
    const double texels_d = (double)s.size_cos_theta * (double)s.size_roughness * (double)s.size_ior;
    if (texels_d > (double)(1 << 26))
        fail(MIPT_ERR_INVALID_ARGUMENT, "bake: at most 2^26 texels");
    if (s.integration_sample_count < 1 || s.integration_sample_count > (1 << 30))
        fail(MIPT_ERR_INVALID_ARGUMENT, "bake: integration_sample_count must be in [1, 2^30]");
    if (s.masking_shadowing != MIPT_GGX_HEIGHT_CORRELATED && s.masking_shadowing != MIPT_GGX_HEIGHT_UNCORRELATED)
        fail(MIPT_ERR_INVALID_ARGUMENT, "bake: unknown masking_shadowing flavour");
    if (!out)
        fail(MIPT_ERR_INVALID_ARGUMENT, "bake: out_device is NULL");
    int32_t ipk, launches, normalizer;
    bake_plan(s, ipk, launches, normalizer);
    const uint32_t texels = (uint32_t)texels_d;

    int32_t lanes = s.lanes_per_texel;
    if (lanes == 0)
    {
        lanes = 1;
        while (lanes < 64 && (uint64_t)texels * (uint64_t)lanes < FILL_THREADS && 2 * lanes <= launches)
            lanes *= 2;
    }
    if (lanes < 1 || lanes > 64 || (lanes & (lanes - 1)) != 0)
        fail(MIPT_ERR_INVALID_ARGUMENT, "bake: lanes_per_texel must be 0 or a power of two up to 64");
    if ((uint64_t)texels * (uint64_t)lanes > (1ull << 31))
        fail(MIPT_ERR_INVALID_ARGUMENT, "bake: texels x lanes_per_texel must stay within 2^31");
    int32_t lane_bits = 0;
    while ((1 << lane_bits) < lanes)
        lane_bits++;
    const int32_t per_lane = (launches + lanes - 1) / lanes; // launches of the busiest lane
    const uint64_t threads = (uint64_t)texels << lane_bits;
    const double samples_per_m = (double)threads * (double)ipk;
    int32_t m_step = (int32_t)std::fmax(1.0, std::floor(SAMPLES_PER_KERNEL / samples_per_m));
    if (m_step > per_lane)
        m_step = per_lane;

    DeviceBuffer partial;
    if (m_step < per_lane)
        partial.reserve(4 * (size_t)threads);
    BakeParams p{};
    p.nx = s.size_cos_theta, p.ny = s.size_roughness, p.nz = s.size_ior;
    p.ipk = ipk, p.launches = launches, p.lane_bits = lane_bits;
    p.normalizer = (float)normalizer;
    p.texels = texels;
    const dim3 grid((unsigned)((threads + 255) / 256));
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, 0));
    int32_t kernel_launches = 0;
    for (int32_t m = 0; m < per_lane; m += m_step)
    {
        p.m_begin = m;
        p.m_end = m + m_step < per_lane ? m + m_step : per_lane;
        p.first = m == 0;
        p.last = p.m_end == per_lane;
        launch_table(s.table, s.masking_shadowing == MIPT_GGX_HEIGHT_CORRELATED, grid, p, partial.as<float>(), out);
        HIP_CHECK(hipGetLastError());
        kernel_launches++;
    }
    HIP_CHECK(hipEventRecord(e1, 0));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
    if (info)
    {
        info->iterations_per_launch = ipk;
        info->launches = launches;
        info->normalizer = normalizer;
        info->lanes_per_texel = lanes;
        info->kernel_launches = kernel_launches;
        info->ms = ms;
        info->samples = (double)texels * (double)launches * (double)ipk;
    }
}

} // namespace mipt
