/*
 * oracle.h -- CPU restatement of the HIPRT-Path-Tracer hot path, used ONLY as
 * a checker.
 *
 * This is test infrastructure. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only to compute the expected
 * answer (or to time the CPU baseline). Nothing in the product links it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root). The code is written from
 * the behaviour, not copied.
 *
 * Floating point: compiled with -ffp-contract=off, no -ffast-math, so every
 * float operation is a single IEEE-754 binary32 operation in source order.
 * The GPU kernels are compiled the same way; this is what makes hit IDs and
 * hit distances bit-identical between the two.
 */
#ifndef MIPT_ORACLE_H
#define MIPT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Ray record: 32 bytes. `last_prim` is the self-intersection filter
 * (FilterFunction.h:37-45); `tmax` is exclusive (hits need t < tmax). */
typedef struct {
    float o[3];
    int32_t last_prim;
    float d[3];
    float tmax;
} orc_ray;

/* Hit record: 16 bytes. Miss = {prim=-1, t=-1, u=0, v=0}. */
typedef struct {
    int32_t prim;
    float t;
    float u;
    float v;
} orc_hit;

/* HIPRTCamera (src/HostDeviceCommon/HIPRTCamera.h:16-27): three row-major
 * 4x4 matrices, position, fov, sensor size, jitter flag. */
typedef struct {
    float inverse_view[4][4];
    float inverse_projection[4][4];
    float view_projection[4][4];
    float position[3];
    float vertical_fov;
    int32_t sensor_width;
    int32_t sensor_height;
    int32_t do_jittering;
} orc_camera;

/* ---- RNG: src/Device/includes/Hash.h:11-19, src/HostDeviceCommon/Xorshift.h:18-66 */
uint32_t orc_wang_hash(uint32_t seed);
uint32_t orc_xorshift32(uint32_t* state);
float orc_rand01(uint32_t* state);
int32_t orc_random_index(uint32_t* state, int32_t n);
uint32_t orc_pixel_seed(uint32_t pixel_index, uint32_t sample_number,
                        uint32_t random_number, int32_t freeze_random);
void orc_xorshift_fill(uint32_t seed, uint32_t* out, int32_t n);
void orc_rand01_fill(uint32_t seed, float* out, int32_t n);

/* ---- Moller-Trumbore: src/Renderer/Triangle.h:20-63 */
int32_t orc_mt_intersect(const float* a, const float* b, const float* c,
                         const orc_ray* ray, float* t, float* u, float* v);

/* ---- Brute force closest / any hit over every triangle (the hit-ID truth).
 * verts: float[3*V], idx: int[3*T]. any_hit != 0 -> hits[i].prim is the
 * lowest primitive index of any valid hit (occlusion only needs prim >= 0). */
void orc_trace_brute(const float* verts, const int32_t* idx, int32_t ntri,
                     const orc_ray* rays, orc_hit* hits, int32_t nrays,
                     int32_t any_hit, int32_t nthreads);

/* ---- Octree of 7-plane k-DOPs: src/Renderer/BVH.h:24-262, BVH.cpp:12-70,
 * BoundingVolume.h:14-131. */
typedef struct orc_octree orc_octree;
orc_octree* orc_octree_build(const float* verts, const int32_t* idx, int32_t ntri,
                             int32_t max_depth, int32_t leaf_max);
void orc_octree_free(orc_octree* tree);
void orc_octree_trace(const orc_octree* tree, const orc_ray* rays, orc_hit* hits,
                      int32_t nrays, int32_t any_hit, int32_t nthreads);
/* stats[0]=nodes, stats[1]=leaves, stats[2]=max depth, stats[3]=tri refs */
void orc_octree_stats(const orc_octree* tree, int64_t* stats);

/* ---- Camera: src/HostDeviceCommon/HIPRTCamera.h:31-52 and the jitter of
 * src/Device/kernels/CameraRays.h:155-172. Writes W*H rays, pix = x + y*W. */
void orc_camera_ray(const orc_camera* cam, float x, float y, int32_t resx, int32_t resy,
                    orc_ray* out);
void orc_camera_rays(const orc_camera* cam, int32_t w, int32_t h, uint32_t sample_number,
                     uint32_t random_number, int32_t freeze_random,
                     orc_ray* out_rays, uint32_t* out_rng_state);

/* ---- Hit attributes, restating src/Device/includes/Intersect.h:222-245
 * (geometric normal = normalize(cross(e1,e2)), inter point = o + t*d,
 * fix_backfacing_normals :142-176) for meshes without vertex normals. */
void orc_hit_attributes(const float* verts, const int32_t* idx, const orc_ray* ray,
                        const orc_hit* hit, float* inter_point, float* geometric_normal,
                        float* shading_normal);

/* ---- Incoherent bounce rays: one cosine-weighted direction about the
 * shading normal of each primary hit (src/Device/includes/Sampling.h:180-201),
 * drawn from the per-pixel xorshift state that the camera pass left behind.
 * Primary misses produce a copy of the primary ray. */
void orc_bounce_rays(const float* verts, const int32_t* idx, const orc_ray* primary,
                     const orc_hit* hits, uint32_t* rng_state, int32_t n, orc_ray* out);

#ifdef __cplusplus
}
#endif
#endif
