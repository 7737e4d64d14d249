// camera.hip -- primary-ray generation and incoherent bounce rays.
//
// Camera rays restate CameraRays (src/Device/kernels/CameraRays.h:155-172):
// per-pixel seed wang_hash((pix+1)*(sample+1)*random_number), two jitter
// draws, then HIPRTCamera::get_camera_ray (src/HostDeviceCommon/HIPRTCamera.h:31-52)
// with matrix_X_point (src/HostDeviceCommon/Math.h:617-634). With
// -ffp-contract=off the rays are bit-identical to the CPU restatement.
//
// Bounce rays: hit point o + t*d, geometric normal normalize(cross(e1, e2))
// (src/Renderer/Triangle.h:54, src/Device/includes/Intersect.h:226-233),
// fix_backfacing_normals (Intersect.h:142-176), then
// cosine_weighted_sample_around_normal_world_space (src/Device/includes/Sampling.h:180-201).
#include <hip/hip_runtime.h>

#include <cstring>

#include "internal.h"
#include "pt_math.h"

namespace mipt {
namespace {

struct CamDev {
    float iv[4][4];
    float ip[4][4];
    int32_t jitter;
};

__device__ __forceinline__ pt::f3 matrix_x_point(const float (*m)[4], pt::f3 p)
{
    float xt = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
    float yt = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
    float zt = m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3];
    float wt = m[3][0] * p.x + m[3][1] * p.y + m[3][2] * p.z + m[3][3];
    float inv_w = 1.0f;
    if (!(wt < 1.0e-10f && wt > -1.0e-10f))
        inv_w = 1.0f / wt;
    return pt::mk(xt * inv_w, yt * inv_w, zt * inv_w);
}

// Ray of pixel (x, y) of a w x h view: seed of its index in the view, two
// jitter draws, get_camera_ray. Writes slot `out` (the pixel's index in the
// buffer: the same in a single view, offset in a batch of views).
__device__ __forceinline__ void camera_ray(const float (*iv)[4], const float (*ip)[4], int32_t jitter, int32_t x,
                                           int32_t y, int32_t w, int32_t h, uint32_t sample_number,
                                           uint32_t random_number, int32_t freeze, float4* __restrict__ rays,
                                           uint32_t* __restrict__ rng_out, uint32_t out)
{
    const uint32_t pix = (uint32_t)x + (uint32_t)y * (uint32_t)w;
    pt::Xorshift32 rng{pt::pixel_seed(pix, sample_number, random_number, freeze != 0)};
    float px = (float)x + 0.5f;
    float py = (float)y + 0.5f;
    if (jitter)
    {
        px += rng() - 0.5f;
        py += rng() - 0.5f;
    }
    const float xn = px / (float)w * 2.0f - 1.0f;
    const float yn = py / (float)h * 2.0f - 1.0f;
    const pt::f3 origin = matrix_x_point(iv, pt::mk(0.0f, 0.0f, 0.0f));
    const pt::f3 p_vs = matrix_x_point(ip, pt::mk(xn, yn, -1.0f));
    const pt::f3 p_ws = matrix_x_point(iv, p_vs);
    const pt::f3 dir = pt::normalize(p_ws - origin);
    rays[2 * (size_t)out] = make_float4(origin.x, origin.y, origin.z, __int_as_float(-1));
    rays[2 * (size_t)out + 1] = make_float4(dir.x, dir.y, dir.z, INFINITY);
    if (rng_out)
        rng_out[out] = rng.state;
}

__global__ void k_camera_rays(CamDev cam, int32_t w, int32_t h, uint32_t sample_number, uint32_t random_number,
                              int32_t freeze, float4* __restrict__ rays, uint32_t* __restrict__ rng_out,
                              const FrameParams* __restrict__ frame)
{
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h)
        return;
    if (frame)
    {
        sample_number = frame->sample_number;
        random_number = frame->rn_camera;
    }
    camera_ray(cam.iv, cam.ip, cam.jitter, x, y, w, h, sample_number, random_number, freeze, rays, rng_out,
               (uint32_t)x + (uint32_t)y * (uint32_t)w);
}

// The rays of a batch of views, one thread per pixel of the batch; the view
// of a pixel by a binary search over the views' first pixels.
__global__ void k_camera_rays_views(const ViewDev* __restrict__ views, int32_t view_count, uint32_t total,
                                    uint32_t sample_number, uint32_t random_number, int32_t freeze,
                                    float4* __restrict__ rays, const FrameParams* __restrict__ frame)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= total)
        return;
    if (frame)
    {
        sample_number = frame->sample_number;
        random_number = frame->rn_camera;
    }
    int32_t lo = 0, hi = view_count - 1;
    while (lo < hi)
    {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (views[mid].start <= pix)
            lo = mid;
        else
            hi = mid - 1;
    }
    const ViewDev& v = views[lo];
    const uint32_t local = pix - v.start;
    const int32_t x = (int32_t)(local % (uint32_t)v.width), y = (int32_t)(local / (uint32_t)v.width);
    camera_ray(v.iv, v.ip, v.jitter, x, y, v.width, v.height, sample_number, random_number, freeze, rays, nullptr,
               pix);
}

__device__ __forceinline__ pt::f3 reflect_ray(pt::f3 d, pt::f3 n)
{
    float k = 2.0f * pt::dot(d, n);
    return k * n - d;
}

__device__ __forceinline__ void fix_backfacing(pt::f3 view, pt::f3& gn, pt::f3& sn)
{
    if (pt::dot(view, gn) < 0.0f)
    {
        gn = -1.0f * gn;
        sn = -1.0f * sn;
    }
    if (pt::dot(view, sn) < 0.0f)
        sn = -1.0f * sn;
    pt::f3 refl = reflect_ray(view, sn);
    if (pt::dot(refl, gn) <= 0.0f)
    {
        const float eps = 0.01f;
        refl = refl - pt::normalize((pt::dot(refl, gn) - eps) * gn);
        sn = pt::normalize(view + refl);
    }
}

__global__ void k_bounce_rays(const float* __restrict__ verts, const int32_t* __restrict__ idx,
                              const float4* __restrict__ primary, const int4* __restrict__ hits,
                              uint32_t* __restrict__ rng_state, float4* __restrict__ out, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float4 ra = primary[2 * (size_t)i], rb = primary[2 * (size_t)i + 1];
    const int4 hit = hits[i];
    if (hit.x < 0)
    {
        out[2 * (size_t)i] = ra;
        out[2 * (size_t)i + 1] = rb;
        return;
    }
    const int32_t prim = hit.x;
    const float t = __int_as_float(hit.y);
    pt::f3 v[3];
    for (int k = 0; k < 3; k++)
    {
        int32_t vi = idx[3 * (int64_t)prim + k];
        v[k] = pt::mk(verts[3 * (int64_t)vi], verts[3 * (int64_t)vi + 1], verts[3 * (int64_t)vi + 2]);
    }
    const pt::f3 o = pt::mk(ra.x, ra.y, ra.z), d = pt::mk(rb.x, rb.y, rb.z);
    const pt::f3 p = o + t * d;
    pt::f3 gn = pt::normalize(pt::normalize(pt::cross(v[1] - v[0], v[2] - v[0])));
    pt::f3 sn = gn;
    fix_backfacing(-1.0f * d, gn, sn);

    pt::Xorshift32 rng{rng_state[i]};
    float r1 = rng();
    float r2 = 2.0f * rng() - 1.0f;
    if (r1 < 1.0e-8f && r2 < -0.999999f && sn.z > 0.999999f)
    {
        r1 += 1.0e-7f;
        r2 += 1.0e-7f;
    }
    const float theta = 6.28318530717958647693f * r1;
    const float s = sqrtf(1.0f - r2 * r2);
    const pt::f3 sp = pt::mk(s * pt::cos_fixed(theta), s * pt::sin_fixed(theta), r2);
    const pt::f3 nd = pt::normalize(sn + sp);
    rng_state[i] = rng.state;
    out[2 * (size_t)i] = make_float4(p.x, p.y, p.z, __int_as_float(prim));
    out[2 * (size_t)i + 1] = make_float4(nd.x, nd.y, nd.z, INFINITY);
}

} // namespace

void camera_rays(Ctx& ctx, const mipt_camera& cam, int32_t w, int32_t h, uint32_t sample_number,
                 uint32_t random_number, int32_t freeze, mipt_ray* rays, uint32_t* rng, hipStream_t stream,
                 const FrameParams* frame)
{
    (void)ctx;
    CamDev c;
    memcpy(c.iv, cam.inverse_view, sizeof(c.iv));
    memcpy(c.ip, cam.inverse_projection, sizeof(c.ip));
    c.jitter = cam.do_jittering;
    dim3 block(8, 8);
    dim3 grid((w + 7) / 8, (h + 7) / 8);
    hipLaunchKernelGGL(k_camera_rays, grid, block, 0, stream, c, w, h, sample_number, random_number, freeze,
                       reinterpret_cast<float4*>(rays), rng, frame);
    HIP_CHECK(hipGetLastError());
}

void camera_rays_views(Ctx& ctx, uint32_t sample_number, uint32_t random_number, int32_t freeze, mipt_ray* rays,
                       hipStream_t stream, const FrameParams* frame)
{
    const RenderState& R = ctx.render;
    const uint32_t total = R.pixels;
    hipLaunchKernelGGL(k_camera_rays_views, dim3((total + 255) / 256), dim3(256), 0, stream,
                       R.view_table.as<ViewDev>(), (int32_t)R.views.size(), total, sample_number, random_number,
                       freeze, reinterpret_cast<float4*>(rays), frame);
    HIP_CHECK(hipGetLastError());
}

void bounce_rays(Ctx& ctx, const mipt_ray* primary, const mipt_hit* hits, uint32_t* rng, mipt_ray* out, uint32_t n,
                 hipStream_t stream)
{
    if (n == 0)
        return;
    if (!ctx.scene_set)
        fail(MIPT_ERR_NO_SCENE, "no scene: call mipt_set_scene first");
    hipLaunchKernelGGL(k_bounce_rays, dim3((n + 255) / 256), dim3(256), 0, stream, ctx.vertices.as<float>(),
                       ctx.indices.as<int32_t>(), reinterpret_cast<const float4*>(primary),
                       reinterpret_cast<const int4*>(hits), rng, reinterpret_cast<float4*>(out), n);
    HIP_CHECK(hipGetLastError());
}

} // namespace mipt
