/*
 * mipt.h -- C ABI of the MI355X path-tracing core (libmipt.so).
 *
 * This is the drop-in boundary for the GPU kernel path of HIPRT-Path-Tracer.
 * Every entry point names the reference interface it replaces (paths relative
 * to the reference repository root). Plain C types only: host pointers are
 * copied, device pointers are caller-owned HIP allocations, `stream` is a
 * hipStream_t passed as void* (NULL = the null stream).
 *
 * Errors: every int-returning call returns MIPT_OK (0) or a negative
 * MIPT_ERR_* code; mipt_last_error() gives the message. No exception crosses
 * the ABI. A context is not thread-safe; work is serialised on the stream
 * passed in.
 */
#ifndef MIPT_H
#define MIPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MIPT_OK 0
#define MIPT_ERR_INVALID_ARGUMENT -1
#define MIPT_ERR_HIP -2
#define MIPT_ERR_NO_SCENE -3
#define MIPT_ERR_NO_BVH -4
#define MIPT_ERR_OUT_OF_MEMORY -5
#define MIPT_ERR_UNSUPPORTED -6

/* ------------------------------------------------------------------ */
/* Ray queries                                                          */
/* ------------------------------------------------------------------ */

/* 32-byte ray. Replaces hiprtRay {origin, minT, direction, maxT} plus the
 * FilterFunctionPayload::last_hit_primitive_index self-intersection filter
 * (src/Device/includes/Intersect.h:38-56, src/Device/functions/FilterFunction.h:37-45).
 * minT is always 0 in the reference; `tmax` is exclusive: a hit needs t < tmax. */
typedef struct mipt_ray {
    float o[3];
    int32_t last_prim; /* primitive the ray leaves, -1 for none */
    float d[3];
    float tmax;        /* +inf for path rays; t_max - 1e-4 for shadow rays (Intersect.h:285) */
} mipt_ray;

/* 16-byte hit. Replaces hiprtHit {primID, t, uv} (normal is recomputed from
 * the triangle, src/Renderer/Triangle.h:54). Miss = {-1, -1.0f, 0, 0}. */
typedef struct mipt_hit {
    int32_t prim;
    float t;
    float u;
    float v;
} mipt_hit;

/* ------------------------------------------------------------------ */
/* Camera: field-for-field mirror of HIPRTCamera                        */
/* (src/HostDeviceCommon/HIPRTCamera.h:16-27), matrices row-major.      */
/* ------------------------------------------------------------------ */
typedef struct mipt_camera {
    float inverse_view[4][4];
    float inverse_projection[4][4];
    float view_projection[4][4];
    float position[3];
    float vertical_fov;
    int32_t sensor_width;
    int32_t sensor_height;
    int32_t do_jittering;
} mipt_camera;

/* ------------------------------------------------------------------ */
/* Materials: mirror of CPUMaterial (src/HostDeviceCommon/Material/     */
/* MaterialCPU.h:213-330). Same names, same defaults (see               */
/* mipt_material_default). Packed on upload the way                     */
/* CPUMaterial::pack_to_GPU (MaterialCPU.h:31-150) packs it.            */
/* ------------------------------------------------------------------ */
typedef struct mipt_material {
    float emission[3];
    float emission_strength;
    float global_emissive_factor;
    int32_t emissive_texture_used;
    float base_color[3];
    float roughness;
    float oren_nayar_sigma;
    float metallic;
    float metallic_F90_falloff_exponent;
    float metallic_F82[3];
    float metallic_F90[3];
    float anisotropy;
    float anisotropy_rotation;
    float second_roughness_weight;
    float second_roughness;
    int32_t do_metallic_energy_compensation;
    float specular;
    float specular_tint;
    float specular_color[3];
    float specular_darkening;
    int32_t do_specular_energy_compensation;
    float coat;
    float coat_medium_absorption[3];
    float coat_medium_thickness;
    float coat_roughness;
    float coat_roughening;
    float coat_darkening;
    float coat_anisotropy;
    float coat_anisotropy_rotation;
    float coat_ior;
    int32_t do_coat_energy_compensation;
    float sheen;
    float sheen_roughness;
    float sheen_color[3];
    float ior;
    float specular_transmission;
    float diffuse_transmission;
    float absorption_at_distance;
    float absorption_color[3];
    float dispersion_scale;
    float dispersion_abbe_number;
    int32_t thin_walled;
    int32_t do_glass_energy_compensation;
    float thin_film;
    float thin_film_ior;
    float thin_film_thickness;
    float thin_film_kappa_3;
    float thin_film_hue_shift_degrees;
    float thin_film_base_ior_override;
    int32_t thin_film_do_ior_override;
    float alpha_opacity;
    int32_t dielectric_priority;
} mipt_material;

/* Fills `m` with the CPUMaterial defaults (MaterialCPU.h:213-330). */
void mipt_material_default(mipt_material* m);

/* ------------------------------------------------------------------ */
/* Scene upload: mirror of struct Scene (src/Scene/SceneParser.h:109-170)*/
/* All pointers are HOST arrays, copied by mipt_set_scene.              */
/* ------------------------------------------------------------------ */
typedef struct mipt_scene_desc {
    const int32_t* triangles_vertex_indices;    /* int[3*triangle_count] */
    int32_t triangle_count;
    const float* vertices_positions;            /* float[3*vertex_count] */
    int32_t vertex_count;
    const uint8_t* has_vertex_normals;          /* uint8[vertex_count] or NULL */
    const float* vertex_normals;                /* float[3*vertex_count] or NULL */
    const int32_t* material_indices;            /* int[triangle_count] or NULL */
    const mipt_material* materials;             /* [material_count] or NULL */
    int32_t material_count;
    const int32_t* emissive_triangles_primitive_indices; /* int[emissive_triangle_count] */
    int32_t emissive_triangle_count;
} mipt_scene_desc;

/* BVH builders (replace hiprtBuildGeometry, src/HIPRT-Orochi/HIPRTScene.h:87-150).
 * LBVH: Morton-ordered Karras hierarchy with SAH-collapsed leaves (fast
 * build, the hiprtBuildFlagBitPreferFastBuild role). SAH_BINNED: top-down
 * binned SAH, 32 bins per axis (the hiprtBuildFlagBitPreferHighQualityBuild
 * role the reference requests). Both give identical hits. */
#define MIPT_BVH_LBVH 0
#define MIPT_BVH_SAH_BINNED 1

typedef struct mipt_bvh_info {
    int32_t builder;
    int32_t triangle_count;
    int32_t node_count;        /* 64-byte nodes written */
    int32_t leaf_count;
    int32_t max_depth;         /* deepest leaf, root = 0 */
    int32_t max_leaf_size;
    int64_t node_bytes;
    int64_t triangle_bytes;
    double build_ms;           /* device time of the build kernels */
    double sah_cost;           /* SAH cost (C_node = 1, C_tri = 1), relative to root area */
} mipt_bvh_info;

/* ------------------------------------------------------------------ */
/* Context                                                              */
/* ------------------------------------------------------------------ */
typedef struct mipt_ctx mipt_ctx;

/* Replaces HIPRTOrochiCtx::init (src/HIPRT-Orochi/HIPRTOrochiCtx.h:40-106). */
int mipt_create(int hip_device, mipt_ctx** out);
void mipt_destroy(mipt_ctx* ctx);
/* Message of the context's last failure; with ctx == NULL, of the calling
 * thread's last mipt_create failure. */
const char* mipt_last_error(const mipt_ctx* ctx);
const char* mipt_version(void);

/* Replaces GPURenderer::set_scene (src/Renderer/GPURenderer.cpp:952-960 ->
 * set_hiprt_scene_from_scene :822-940): uploads geometry, materials
 * (packed), emissive list. Does not build the BVH. */
int mipt_set_scene(mipt_ctx* ctx, const mipt_scene_desc* scene);

/* Replace GPURenderer::update_all_materials / update_one_material
 * (src/Renderer/GPURenderer.cpp:1022-1054): new CPU materials for the scene
 * (count = the scene's material count), packed and uploaded without a BVH
 * rebuild. As in the reference, the emissive triangle list and the power
 * alias table built from it stay as they were. Waits for the device before
 * overwriting the material records. */
int mipt_update_materials(mipt_ctx* ctx, const mipt_material* materials, int32_t count);
int mipt_update_material(mipt_ctx* ctx, int32_t index, const mipt_material* material);

/* Replaces HIPRTGeometry::build_bvh (HIPRTScene.h:87-150). Synchronous. */
int mipt_build_bvh(mipt_ctx* ctx, int32_t builder);
int mipt_get_bvh_info(const mipt_ctx* ctx, mipt_bvh_info* out);
/* Copies the BVH to the host: node_count 64-byte nodes (root = node 0) and
 * triangle_count 48-byte triangle records in leaf order (the layouts of
 * csrc/layout.h); either pointer may be NULL. For inspection and tests: the
 * reference's HIPRT keeps its BVH opaque, so this has no counterpart there. */
int mipt_read_bvh(const mipt_ctx* ctx, void* host_nodes, void* host_triangles);

/* Closest hit over the scene BVH: replaces hiprtGeomTraversalClosest
 * (used by trace_main_path_ray, src/Device/includes/Intersect.h:200-270, and
 * evaluate_bsdf_light_sample_ray :544-643). Ties in t go to the smaller
 * primitive index. d_rays/d_hits are device pointers of n records. */
int mipt_trace_closest(mipt_ctx* ctx, const mipt_ray* d_rays, mipt_hit* d_hits, uint32_t n,
                       void* stream);

/* Any hit: replaces hiprtGeomTraversalAnyHit in evaluate_shadow_ray_occluded
 * (Intersect.h:276-324). d_occluded[i] = 1 if some hit has t < tmax. */
int mipt_trace_any(mipt_ctx* ctx, const mipt_ray* d_rays, uint8_t* d_occluded, uint32_t n,
                   void* stream);

/* Closest hit that allocates nothing: the result of mipt_trace_closest,
 * with the traversal's work counter and spill stack in a static device
 * scratch array of MIPT_TRACE_NOALLOC_SCRATCH_BYTES per GPU (held from the
 * library's load until the process ends) instead of memory the context
 * allocates. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc). It calls no hipMalloc/hipFree.
 * When the tree's spill columns for the whole persistent grid exceed the
 * scratch, fewer blocks trace the rays; a BVH too deep for one block's
 * spill stack is refused (MIPT_ERR_INVALID_ARGUMENT). Every call on one GPU
 * shares that GPU's scratch: calls that overlap in time on one GPU (two
 * streams, two host threads) race and must be serialised by the caller.
 * The BVH must be built (mipt_build_bvh). */
#define MIPT_TRACE_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
int mipt_trace_closest_noalloc(mipt_ctx* ctx, const mipt_ray* d_rays, mipt_hit* d_hits, uint32_t n,
                               void* stream);

/* Instrumented closest hit: same result as mipt_trace_closest and, per ray,
 * d_counts[2*i] = 64-byte nodes fetched, d_counts[2*i+1] = triangles tested.
 * Used for the bytes-per-ray accounting; slower than the plain kernel. */
int mipt_trace_closest_counted(mipt_ctx* ctx, const mipt_ray* d_rays, mipt_hit* d_hits,
                               uint32_t* d_counts, uint32_t n, void* stream);

/* Order in which the path pass shades the live paths of a bounce. The
 * reference's megakernel shades one pixel per thread (Megakernel.h:23-144);
 * the wavefront passes shade a list of live paths, which is either kept in
 * the order the paths reached their vertex (LIVE) or sorted by the lobe
 * class of the material at the vertex and by whether the path carries a
 * volume state (MATERIAL: the paths of a wave then take the same BSDF
 * branches). AUTO (the default) sorts when the scene has transmissive
 * materials (glass, diffuse transmission), where paths inside media and at
 * opaque surfaces would share waves. Results do not depend on the order. */
#define MIPT_SHADE_ORDER_LIVE 0
#define MIPT_SHADE_ORDER_MATERIAL 1
#define MIPT_SHADE_ORDER_AUTO 2
int mipt_set_shading_order(mipt_ctx* ctx, int32_t order);

/* ------------------------------------------------------------------ */
/* Batched camera views (many views of uneven resolution rendered as    */
/* one wavefront: each pass kernel and each traversal launch covers the */
/* pixels of every view). No counterpart in the reference, which        */
/* renders one camera.                                                  */
/* ------------------------------------------------------------------ */
typedef struct mipt_view {
    mipt_camera camera; /* HIPRTCamera of the view (sensor size unused) */
    int32_t width, height;
} mipt_view;

/* Switches the context to `count` views: the G-buffer, path state and
 * framebuffer hold the pixels of every view back to back (view v's pixel
 * x + y * width_v at the offset of the views before it), so
 * mipt_read_framebuffer returns them in that order. Each pixel draws the
 * random numbers a render of its view alone would (the seed uses the
 * pixel's index in its view), so a batch equals its views rendered one by
 * one, bit for bit. The render settings' width and height are replaced by
 * (total pixels, 1); the plain path tracer only (MIS light + BSDF, no
 * ReSTIR DI, no radiance cache). count = 0 returns to one view (call
 * mipt_set_render_settings again). */
int mipt_set_views(mipt_ctx* ctx, const mipt_view* views, int32_t count);

/* Seeded view batch (configuration 3): `count` views of the config-1 scene
 * with heights log-uniform in [128, 1024] px, aspect (width / height)
 * uniform in [1, 16/9], cameras inside the box looking at its back wall
 * (PCG32, state 0x2026120200000000 + 3, stream 3). */
int mipt_views_generate(int32_t count, mipt_view* out);

/* Graph replay of plain samples: with enable = 1, the first
 * mipt_render_sample with sample_number > 0 after any change of scene,
 * settings, camera, views or launch configuration records the sample's
 * kernels (camera pass + path pass) into a hipGraph; later samples replay it,
 * the per-sample values (sample number, host random numbers, consistent
 * tau) written to device memory by one small kernel before each replay.
 * Results equal the eager launches bit for bit. Sample 0 always runs
 * eagerly (it resets the accumulation). Plain path tracing only. */
int mipt_set_graph_capture(mipt_ctx* ctx, int32_t enable);
/* Replays since the last capture (0 when the last sample ran eagerly). */
int mipt_graph_replays(const mipt_ctx* ctx, uint32_t* out);

/* Traversal launch configuration (the autotuned knobs). block_size in
 * {64,128,256}; lds_stack in {8,16,24,32}. 0 keeps the current value. */
int mipt_set_trace_config(mipt_ctx* ctx, int32_t block_size, int32_t lds_stack);
/* Queue tile: rays a wave of the persistent traversal takes per
 * work-counter atomic, in batches of 64 (1, 2 or 4). */
int mipt_set_queue_tile(mipt_ctx* ctx, int32_t tile);

/* Autotuner of the traversal launch (no counterpart in the reference,
 * whose HIPRT kernels are compiled with fixed block and stack sizes,
 * MegaKernelRenderPass.cpp:22-23): times `repeats` closest-hit launches
 * over the caller's n device rays for every block size (64, 128, 256), LDS
 * stack depth (8, 16, 24, 32) and queue tile (1, 2, 4), and sets the
 * fastest on the context. table_ms (36 doubles, [block][stack][tile], may be
 * NULL) receives every mean launch time. Synchronous. */
typedef struct mipt_trace_tuning {
    int32_t block_size, lds_stack, queue_tile;
    int32_t configurations; /* timed */
    double best_ms;         /* mean launch time of the chosen configuration */
    double default_ms;      /* of 128 threads, 16 LDS entries, tile 1 */
} mipt_trace_tuning;
int mipt_autotune_trace(mipt_ctx* ctx, const mipt_ray* d_rays, uint32_t n, int32_t repeats, void* stream,
                        mipt_trace_tuning* out, double* table_ms);

/* ------------------------------------------------------------------ */
/* Camera rays (the ray-generation half of CameraRays,                  */
/* src/Device/kernels/CameraRays.h:155-172 + HIPRTCamera.h:31-52).       */
/* Writes w*h rays, pix = x + y*w, y = 0 the bottom row, and the        */
/* per-pixel xorshift32 state after the jitter draws.                   */
/* ------------------------------------------------------------------ */
int mipt_camera_rays(mipt_ctx* ctx, const mipt_camera* cam, int32_t w, int32_t h,
                     uint32_t sample_number, uint32_t random_number, int32_t freeze_random,
                     mipt_ray* d_rays, uint32_t* d_rng_state, void* stream);

/* Incoherent bounce rays: from each primary hit, one cosine-weighted
 * direction about the (back-face fixed) shading normal
 * (src/Device/includes/Sampling.h:180-201, Intersect.h:142-176), drawn from
 * d_rng_state (updated). Misses copy the primary ray. */
int mipt_bounce_rays(mipt_ctx* ctx, const mipt_ray* d_primary, const mipt_hit* d_hits,
                     uint32_t* d_rng_state, mipt_ray* d_out, uint32_t n, void* stream);


/* ------------------------------------------------------------------ */
/* Path-tracing passes: the per-pass launch surface of the renderer.    */
/* ------------------------------------------------------------------ */

/* Render settings: the subset of HIPRTRenderSettings
 * (src/HostDeviceCommon/RenderSettings.h:26-278), MicrofacetRegularizationSettings
 * and the kernel options the passes read. Defaults: mipt_render_settings_default. */
typedef struct mipt_render_settings {
    int32_t width, height;              /* render_resolution */
    int32_t nb_bounces;                 /* RenderSettings.h:101 */
    int32_t do_russian_roulette;        /* :103, default 1 */
    int32_t russian_roulette_min_depth; /* :107, 0 as the reference initialises it */
    float russian_roulette_throughput_clamp; /* :117, default 10 */
    int32_t freeze_random;              /* :123 */
    int32_t enable_direct;              /* :46, default 1 */
    float microfacet_regularization_tau_0;  /* MicrofacetRegularizationSettings.h, 30 */
    float microfacet_regularization_min_roughness; /* 0 */
    int32_t do_microfacet_regularization;   /* PrincipledBSDFKernelOptions.h:171, 1 */
    int32_t accumulate_fp64;            /* keep an FP64 mirror of the framebuffer sum */
    /* DirectLightSamplingStrategy / ...BaseStrategy (kernel options in the
     * reference, DirectLightSamplingOptions.h:20-35): MIPT_LSS_* below.
     * Defaults: MIS light + BSDF, uniform. */
    int32_t direct_light_sampling_strategy;
    int32_t direct_light_sampling_base_strategy;
    /* 1: query the learned radiance cache (see mipt_set_radiance_cache); not
     * a reference setting. Default 0. */
    int32_t radiance_cache;
    /* RenderSettings.h:203 (default 1): the shading normal of a hit on a
     * triangle whose first vertex has a vertex normal is the normalised
     * barycentric interpolation of its three vertex normals
     * (get_shading_normal, Intersect.h:114-133; normal maps need textures,
     * which this build does not upload); 0 keeps the geometric normal. */
    int32_t do_normal_mapping;
    /* RenderSettings::accumulate (RenderSettings.h:71-73), default 1. 0 is
     * the interactive mode: every frame is sample 0 whatever sample_number
     * the call passes (MegaKernelRenderPass.cpp:42-43: the framebuffer is
     * overwritten, the per-pixel seeds use sample 0), the G-buffers and
     * ReSTIR DI reservoirs are not reset between frames, mipt_render_sample
     * draws its random numbers from a running frame count, and ReSTIR DI
     * temporal reuse reprojects each pixel's primary hit into the previous
     * frame's camera to find its temporal neighbour
     * (find_temporal_neighbor_index, UtilsTemporal.h:37-99). */
    int32_t accumulate;
    /* RISSettings (RIS/RISSettings.h:9-16, RenderSettings.h:206): candidates
     * of resampled importance sampling of the lights (MIPT_LSS_RIS_BSDF_AND_LIGHT,
     * and ReSTIR DI's later bounces), defaults 4 light and 1 BSDF candidate
     * (0 or 1 BSDF candidate in this build). */
    int32_t ris_number_of_light_candidates;
    int32_t ris_number_of_bsdf_candidates;
} mipt_render_settings;

#define MIPT_LSS_MIS_LIGHT_BSDF 3 /* LSS_MIS_LIGHT_BSDF: one light sample + one BSDF sample, balance heuristic */
#define MIPT_LSS_RIS_BSDF_AND_LIGHT 4 /* LSS_RIS_BSDF_AND_LIGHT: RIS over light + BSDF candidates (sample_lights_RIS, RIS/RIS.h:252) */
#define MIPT_LSS_RESTIR_DI 5      /* LSS_RESTIR_DI: ReSTIR DI reservoir at bounce 0, later_bounces_sampling_strategy after */
#define MIPT_LSS_BASE_UNIFORM 0   /* LSS_BASE_UNIFORM: emissive triangle picked uniformly */
#define MIPT_LSS_BASE_POWER 1     /* LSS_BASE_POWER: picked by emitted power (alias table) */

void mipt_render_settings_default(mipt_render_settings* s);

/* ReSTIR DI settings: ReSTIRDISettings (src/HostDeviceCommon/ReSTIR/ReSTIRDISettings.h:
 * 17-163) and its ReSTIRCommonSettings (ReSTIRCommonSettings.h). The kernel
 * options of ReSTIRDIOptions.h are fixed at the reference defaults:
 * pairwise MIS bias correction with visibility, visibility reuse, no
 * visibility in the initial or spatial target functions, presampled lights
 * drawn by power (when do_light_presampling is on). Later bounces use
 * later_bounces_sampling_strategy (RIS by default, as the reference). Defaults:
 * mipt_restir_di_settings_default (the reference's). */
typedef struct mipt_restir_di_settings {
    int32_t number_of_initial_light_candidates; /* 4 */
    int32_t number_of_initial_bsdf_candidates;  /* 1 */
    int32_t do_temporal_reuse_pass;             /* 1 */
    int32_t do_spatial_reuse_pass;              /* 1 */
    int32_t spatial_number_of_passes;           /* 1 */
    int32_t spatial_reuse_radius;               /* 16 px */
    int32_t spatial_reuse_neighbor_count;       /* 5, at most 32 */
    int32_t m_cap;                              /* 3; 0 = no cap */
    int32_t use_confidence_weights;             /* 1 */
    int32_t use_normal_similarity_heuristic;    /* 1 */
    float normal_similarity_angle_precomp;      /* cos(37.5 deg) = 0.906307787 */
    int32_t use_plane_distance_heuristic;       /* 1 */
    float plane_distance_threshold;             /* 0.1 */
    int32_t reject_using_geometric_normals;     /* 1 */
    int32_t do_final_shading_visibility;        /* 1 */
    /* Not a reference setting, default 0 (the reference's arithmetic): a BSDF
     * candidate that hits the back of an emissive triangle has an area pdf of
     * 0 and the reference's weight mis * target / pdf is 0/0 = NaN
     * (InitialCandidates.h:402-410), which poisons the reservoir and then its
     * spatial and temporal neighbours. 1 gives such a candidate weight 0. */
    int32_t guard_backfacing_bsdf_candidates;
    /* ReSTIRDISettings::do_fused_spatiotemporal (ReSTIRDISettings.h:141),
     * default 0: temporal reuse and the first spatial pass in one pass
     * (MIPT_PASS_RESTIR_DI_SPATIOTEMPORAL_REUSE). */
    int32_t do_fused_spatiotemporal;
    /* Kernel option ReSTIR_DI_DoLightPresampling (ReSTIRDIOptions.h:151),
     * default 0: light candidates are drawn from subsets of lights presampled
     * by power each frame (MIPT_PASS_RESTIR_DI_LIGHTS_PRESAMPLING), all pixels
     * of a tile_size x tile_size tile from one random subset
     * (ReSTIRDILightPresamplingSettings, ReSTIRDISettings.h:45-56:
     * 128 subsets of 1024 lights, tiles of 8). */
    int32_t do_light_presampling;
    int32_t light_presampling_number_of_subsets;
    int32_t light_presampling_subset_size;
    int32_t light_presampling_tile_size;
    /* ReSTIRDIInitialCandidatesSettings::envmap_candidate_probability
     * (ReSTIRDISettings.h:22), default 0.5: with an envmap, the probability
     * that a light candidate (or a presampled light) is an envmap sample; 1
     * when the scene has no emissive triangle. */
    float envmap_candidate_probability;
    /* ReSTIRCommonTemporalPassSettings (ReSTIRCommonSettings.h:14-34,
     * defaults ReSTIRDISettings.h:65-69), read when render_settings.accumulate
     * is 0: the reprojected pixel is taken if it passes the similarity
     * heuristics against the previous frame's G-buffer, else up to
     * max_neighbor_search_count (8) random pixels within
     * neighbor_search_radius (4) px of it. use_permutation_sampling (0)
     * shuffles the reprojected pixel within 4 x 4 blocks by
     * permutation_sampling_random_bits, which the reference draws from its
     * host stream each frame (ReSTIRDIRenderPass.cpp:470, :518);
     * mipt_render_sample sets it the same way, mipt_launch_pass callers set
     * it here before the temporal or fused pass. */
    int32_t use_permutation_sampling;
    int32_t permutation_sampling_random_bits;
    int32_t max_neighbor_search_count;
    int32_t neighbor_search_radius;
    /* Kernel option ReSTIR_DI_LaterBouncesSamplingStrategy (ReSTIRDIOptions.h:
     * 124-140): the direct light of the bounces after the first,
     * MIPT_RESTIR_DI_LATER_BOUNCES_RIS_BSDF_AND_LIGHT (the default: RIS with
     * render_settings.ris_* candidates) or ..._MIS_LIGHT_BSDF. */
    int32_t later_bounces_sampling_strategy;
} mipt_restir_di_settings;

#define MIPT_RESTIR_DI_LATER_BOUNCES_MIS_LIGHT_BSDF 2    /* RESTIR_DI_LATER_BOUNCES_MIS_LIGHT_BSDF */
#define MIPT_RESTIR_DI_LATER_BOUNCES_RIS_BSDF_AND_LIGHT 3 /* RESTIR_DI_LATER_BOUNCES_RIS_BSDF_AND_LIGHT */

void mipt_restir_di_settings_default(mipt_restir_di_settings* s);
/* Takes effect at the next pass; allocates the reservoir buffers when the
 * strategy is MIPT_LSS_RESTIR_DI. */
int mipt_set_restir_di_settings(mipt_ctx* ctx, const mipt_restir_di_settings* s);

/* Allocates the G-buffer, path state and framebuffer for width x height and
 * packs the scene's materials (replaces GPURenderer::resize + the render-data
 * wiring of update_render_data, GPURenderer.cpp:441, :769-820). Every lobe
 * of the layered Principled BSDF is implemented (coat, sheen, the two metal
 * lobes, dielectric specular, diffuse, glass with dispersion, diffuse
 * transmission, thin film); a dispersive glass is refused
 * (MIPT_ERR_UNSUPPORTED) only under ReSTIR DI, at the pass. */
int mipt_set_render_settings(mipt_ctx* ctx, const mipt_render_settings* s);

/* Replaces GPURenderer::set_camera (GPURenderer.cpp:1102). */
int mipt_set_camera(mipt_ctx* ctx, const mipt_camera* cam);

#define MIPT_PASS_CAMERA_RAYS 0 /* CameraRays (CameraRays.h:83): G-buffer fill */
#define MIPT_PASS_PATH_TRACE 1  /* MegaKernel (Megakernel.h:23): 1 spp, accumulate */
#define MIPT_PASS_RESTIR_DI_INITIAL_CANDIDATES 2 /* ReSTIR_DI_InitialCandidates (DI/InitialCandidates.h:478) */
#define MIPT_PASS_RESTIR_DI_TEMPORAL_REUSE 3     /* ReSTIR_DI_TemporalReuse (DI/TemporalReuse.h:50) */
/* ReSTIR_DI_SpatialReuse (DI/SpatialReuse.h:43) as spatial pass number i
 * (common_spatial_pass.spatial_pass_index), i in [0, spatial_number_of_passes). */
#define MIPT_PASS_RESTIR_DI_SPATIAL_REUSE(i) (16 + (i))
/* ReSTIR_DI_LightsPresampling (DI/LightsPresampling.h:74): the presampled
 * light subsets, with the camera pass's random number (the reference launches
 * it before drawing the initial candidates' number). */
#define MIPT_PASS_RESTIR_DI_LIGHTS_PRESAMPLING 6
/* ReSTIR_DI_SpatiotemporalReuse (DI/FusedSpatiotemporalReuse.h:152):
 * temporal reuse and spatial pass 0 in one pass; further spatial passes
 * (MIPT_PASS_RESTIR_DI_SPATIAL_REUSE(i), i >= 1) follow it. Its random
 * number is the reference's third draw of the pass (after the temporal
 * pass's number and permutation bits, ReSTIRDIRenderPass.cpp:515-604). */
#define MIPT_PASS_RESTIR_DI_SPATIOTEMPORAL_REUSE 7

/* One pass with the given sample_number (render_settings.sample_number) and
 * random_number (render_data.random_number). Asynchronous on `stream`.
 * sample_number 0 resets the accumulation (need_to_reset): the camera pass
 * then clears the G-buffers and the ReSTIR DI reservoirs. The ReSTIR DI
 * passes route their reservoir buffers as ReSTIRDIRenderPass does
 * (src/Renderer/RenderPasses/ReSTIRDIRenderPass.cpp:442-677); the last
 * ReSTIR pass of a frame selects the reservoirs the path pass shades. */
int mipt_launch_pass(mipt_ctx* ctx, int32_t pass, uint32_t sample_number, uint32_t random_number, void* stream);

/* One full sample with the host random stream (xorshift32 seeded with 42)
 * drawn once per pass in the order the GPU renderer draws it: CAMERA_RAYS
 * (FillGBufferRenderPass.cpp:91); with ReSTIR DI, INITIAL_CANDIDATES (one
 * draw, ReSTIRDIRenderPass.cpp:449), TEMPORAL_REUSE (two draws, the second
 * being the unused permutation-sampling bits, :469-470) and each spatial pass
 * (one draw, :532); then PATH_TRACE (MegaKernelRenderPass.cpp:53). Sample s
 * uses draws [s*D, (s+1)*D), D = mipt_draws_per_sample(ctx). */
int mipt_render_sample(mipt_ctx* ctx, uint32_t sample_number, void* stream);
int32_t mipt_draws_per_sample(const mipt_ctx* ctx);

/* Power alias table of the emissive triangles (GPURenderer.cpp:221-295 with
 * Utils::compute_alias_table, src/Utils/Utils.cpp:63-131), built when the
 * base strategy is POWER. Host copies: probas/alias of emissive_triangle_count
 * entries (either may be NULL), *sum = sum of the powers. */
int mipt_get_power_alias_table(mipt_ctx* ctx, float* probas, int32_t* alias, float* sum);

/* Vose alias table of n non-negative weights, normalised in FP64
 * (Utils::compute_alias_table, Utils.cpp:63-131). Host utility. */
int mipt_compute_alias_table(const float* weights, int32_t n, float total_sum, float* out_probas, int32_t* out_alias);

/* ------------------------------------------------------------------ */
/* Ambient light: WorldSettings (src/HostDeviceCommon/WorldSettings.h: */
/* 14-53). A ray that leaves the scene adds the uniform colour at any  */
/* bounce; an environment map is seen by camera rays and otherwise     */
/* importance sampled at every vertex (alias table over the texels'    */
/* luminance, Envmap.h:77-133) with BSDF MIS (Envmap.h:168-255).       */
/* ------------------------------------------------------------------ */
#define MIPT_AMBIENT_NONE 0    /* AmbientLightType::NONE */
#define MIPT_AMBIENT_UNIFORM 1 /* AmbientLightType::UNIFORM (the default, black) */
#define MIPT_AMBIENT_ENVMAP 2  /* AmbientLightType::ENVMAP (needs mipt_set_envmap) */

typedef struct mipt_world_settings {
    int32_t ambient_light_type;                /* :23, default MIPT_AMBIENT_UNIFORM */
    float uniform_light_color[3];              /* :24, default 0 */
    float envmap_intensity;                    /* :30, default 1 */
    /* :35, default 0: camera rays see the map without the intensity */
    int32_t envmap_scale_background_intensity;
    /* :51-52: float3x3 as [column][row] (glm layout), default identity;
     * world_to_envmap is the inverse the caller computed
     * (RendererEnvmap::do_animation, RendererEnvmap.cpp:127-163). */
    float envmap_to_world[3][3];
    float world_to_envmap[3][3];
} mipt_world_settings;

void mipt_world_settings_default(mipt_world_settings* w);
/* Takes effect at the next pass. MIPT_AMBIENT_ENVMAP needs an envmap set
 * with mipt_set_envmap; ReSTIR DI then samples the envmap among its
 * candidates (mipt_restir_di_settings.envmap_candidate_probability), and the
 * radiance cache learns the envmap's light like any other. */
int mipt_set_world_settings(mipt_ctx* ctx, const mipt_world_settings* w);
int mipt_get_world_settings(const mipt_ctx* ctx, mipt_world_settings* out);

/* GPURenderer::set_envmap (GPURenderer.cpp:962-998) for an Image32Bit:
 * `pixels` holds width x height texels of `channels` floats (3 or 4; only
 * the first three are read), row y at v = y / height, the first row looking
 * down (-Y). The texels are packed to R9G9B9E5 on the GPU
 * (RGBE9995Envmap::pack_from, EnvmapRGBE9995.h:21-44); the alias table over
 * the luminance of the float texels and their sum are built on the host
 * (Image32Bit::compute_alias_table, Image.cpp:652-672). Sets the ambient
 * light type to MIPT_AMBIENT_ENVMAP; an empty image (width or height 0)
 * turns an ENVMAP ambient light to UNIFORM, as the reference does. */
int mipt_set_envmap(mipt_ctx* ctx, const float* pixels, int32_t width, int32_t height, int32_t channels);
/* Host copies of the envmap's alias table (width x height entries each,
 * either may be NULL), its luminance sum, and the packed texels. */
int mipt_get_envmap_tables(const mipt_ctx* ctx, float* probas, int32_t* alias, float* luminance_sum,
                           uint32_t* packed_texels);

/* Procedural HDR environment map (no counterpart in the reference, which
 * loads .hdr/.exr files): a sky gradient over a dim ground, each texel
 * scaled by a lognormal factor (sigma 0.35) from PCG32 stream 5, and a sun
 * disk of 1.5 degrees at 35 degrees elevation; width x height x 3 floats in
 * mipt_set_envmap's layout. 4096 x 2048 is the full-stack configuration. */
int mipt_envmap_generate(int32_t width, int32_t height, float* out_rgb);

/* ------------------------------------------------------------------ */
/* Energy-compensation LUT baker: GPUBaker (src/Renderer/Baker/        */
/* GPUBaker.cpp:37-93) and its kernels (src/Device/kernels/Baking/).   */
/* A table holds the Monte Carlo directional albedo E(cos_theta_o,     */
/* roughness[, F0]) of a single-scattering GGX BSDF, isotropic:        */
/*   cos_theta_o = max(1e-3, x / (size_cos_theta - 1)) (^2.5 for the   */
/*     glossy dielectric, glass and Fresnel tables);                   */
/*   roughness = max(1e-4, y / (size_roughness - 1));                  */
/*   F0 = (z / (size_ior - 1))^4, relative IOR (1 + sqrt F0) /          */
/*     (1 - sqrt F0) with F0 clamped to [0, 0.99].                      */
/* Samples per texel: launches x iterations_per_launch as the          */
/* reference's host loop runs them (GPUBakerKernel.cpp:77-92, 1e8      */
/* texel-samples a launch), launch i seeding xorshift32 with           */
/* wang_hash(texel + 1) * i, so a bake draws the reference's samples.  */
/* ------------------------------------------------------------------ */
#define MIPT_LUT_GGX_CONDUCTOR 0       /* GGXConductorDirectionalAlbedo.h, 128 x 128 */
#define MIPT_LUT_GLOSSY_DIELECTRIC 1   /* GlossyDielectricDirectionalAlbedo.h, 128 x 64 x 128 */
#define MIPT_LUT_GGX_GLASS 2           /* GGXGlassDirectionalAlbedo.h entering, 256 x 16 x 128 */
#define MIPT_LUT_GGX_GLASS_INVERSE 3   /* the same, exiting (relative IOR inverted) */
#define MIPT_LUT_GGX_THIN_GLASS 4      /* GGXThinGlassDirectionalAlbedo.h, 32 x 32 x 96 */
#define MIPT_LUT_GGX_FRESNEL 5         /* GGXFresnelDirectionalAlbedo.h, 256 x 256 x 256 (not read by the renderer) */

#define MIPT_GGX_HEIGHT_CORRELATED 0   /* GGXMaskingShadowingFlavor (BSDFsData.h:20-24) */
#define MIPT_GGX_HEIGHT_UNCORRELATED 1 /* the renderer's flavour */

/* The *DirectionalAlbedoSettings structs (src/Renderer/Baker/). */
typedef struct mipt_bake_settings {
    int32_t table;                    /* MIPT_LUT_* */
    int32_t size_cos_theta;           /* texture_size_cos_theta(_o) */
    int32_t size_roughness;           /* texture_size_roughness */
    int32_t size_ior;                 /* texture_size_ior; 1 for the conductor table */
    int32_t integration_sample_count; /* 262144 conductor, 131072 glossy, 65536 others */
    int32_t masking_shadowing;        /* default MIPT_GGX_HEIGHT_UNCORRELATED */
    /* partial sums per texel, 1, 2, 4, ..., 64: launch i goes to partial
     * (i - 1) % lanes, the partials add pairwise ((p0 + p1) + (p2 + p3)) + ...
     * 1 (the default) is the reference's summation order, whose rounding
     * drift (~0.1 %) is in the shipped tables; more lanes run small tables
     * faster and round less; 0 picks lanes for occupancy. */
    int32_t lanes_per_texel;
} mipt_bake_settings;

typedef struct mipt_bake_info {
    int32_t iterations_per_launch; /* samples per texel in one reference launch */
    int32_t launches;              /* reference launches (seeds 1 .. launches) */
    int32_t normalizer;            /* nb_samples dividing each sample (the Fresnel kernel's own count) */
    int32_t lanes_per_texel;       /* partial sums used */
    int32_t kernel_launches;       /* GPU launches the bake took */
    float ms;                      /* GPU time of the bake (hipEvents) */
    double samples;                /* samples drawn, texels x launches x iterations_per_launch */
} mipt_bake_info;

/* Defaults of a table: the reference's settings and the renderer's size. */
int mipt_bake_settings_default(int32_t table, mipt_bake_settings* out);
/* Bakes a table into `out_device` (device memory, size_cos_theta *
 * size_roughness * size_ior floats) in the renderer's layout
 * [z][size_roughness - 1 - y][x]: the reference writes rows in y order and
 * reads them back flipped (GPUBakerKernel.cpp:102-121, Image32Bit::
 * read_image_hdr with flipY). A sample whose value is NaN is skipped like a
 * zero-pdf one. Synchronous; info may be NULL. */
int mipt_bake_lut(mipt_ctx* ctx, const mipt_bake_settings* settings, float* out_device, mipt_bake_info* info);
/* Replaces one of the renderer's tables (MIPT_LUT_GGX_CONDUCTOR ..
 * MIPT_LUT_GGX_THIN_GLASS, the conductor's flavour the uncorrelated one)
 * from host memory in the same layout, for the passes from then on
 * (CPURenderer::setup_brdfs_data, CPURenderer.cpp:155-196, loads the baked
 * files); the sizes must be the renderer's. values NULL restores the table
 * the library ships. */
int mipt_set_lut(mipt_ctx* ctx, int32_t table, const float* values, int32_t size_cos_theta, int32_t size_roughness,
                 int32_t size_ior);
/* Host copy of a table the library ships (the reference's baked .hdr files,
 * RGBE-decoded), in the renderer's layout: MIPT_LUT_GGX_CONDUCTOR with
 * masking_shadowing either flavour, the other tables uncorrelated. */
int mipt_get_shipped_lut(int32_t table, int32_t masking_shadowing, float* out, int32_t count);

/* ------------------------------------------------------------------ */
/* Learned radiance cache. No counterpart in the reference: a neural   */
/* radiance cache trained online from the tracer's own paths. The      */
/* position goes through a multiresolution hash grid (FP16 features,   */
/* Instant-NGP hashing); direction and normal (octahedral) and         */
/* roughness through one-blob encodings; albedo as is. A 6-matrix ReLU */
/* MLP of width 64 with BF16 weights runs on the MFMA units (FP32      */
/* accumulation); relative L2 loss, Adam.                               */
/* ------------------------------------------------------------------ */
typedef struct mipt_radiance_cache_settings {
    int32_t grid_levels;          /* 16 (at most 16) */
    int32_t grid_log2_table;      /* 19: 2^19 entries x 2 features per level */
    int32_t grid_base_resolution; /* 16 */
    float grid_per_level_scale;   /* 1.3819: 2048 cells across the scene at level 15 */
    int32_t batch_size;           /* 65536 training records per step at most */
    float learning_rate;          /* 1e-2 */
    float adam_beta1;             /* 0.9 */
    float adam_beta2;             /* 0.99 */
    float adam_eps;               /* 1e-8 */
    uint32_t seed;                /* parameter initialisation stream */
    int32_t query_bounce;         /* 1: paths stop at this bounce with the cache's radiance */
    int32_t train_every;          /* 64: one path in train_every is a training path */
    int32_t train_extension;      /* 3: training paths trace this many bounces past query_bounce */
    int32_t warmup_train_every;   /* 16: one path in N runs (as a training path) in a warm-up frame */
} mipt_radiance_cache_settings;

void mipt_radiance_cache_settings_default(mipt_radiance_cache_settings* s);
/* (Re)initialises the cache's parameters (Glorot-uniform weights, grid
 * features U(-1e-4, 1e-4), from `seed`) and optimiser state. Needs a scene
 * (the grid spans the scene's bounding cube). */
int mipt_set_radiance_cache(mipt_ctx* ctx, const mipt_radiance_cache_settings* s);

/* Query record: 16 floats = position[3], outgoing direction[3] (towards the
 * viewer), shading normal[3], roughness, albedo[3], 3 unused. Output: 4
 * floats per row (RGB radiance, 0). Device pointers. */
int mipt_radiance_cache_infer(mipt_ctx* ctx, const float* d_queries, uint32_t n, float* d_out, void* stream);
/* The network's input rows for n query records: 64 bfloat16 values per row
 * (grid features level by level, then the one-blob and albedo inputs, a
 * constant 1 and zero padding). Device pointers. */
int mipt_radiance_cache_encode(mipt_ctx* ctx, const float* d_queries, uint32_t n, uint16_t* d_out, void* stream);
/* One optimiser step on n records (targets: 4 floats per row, RGB + unused).
 * *out_loss (host, may be NULL) receives the batch's relative L2 loss;
 * synchronous when out_loss is given. */
int mipt_radiance_cache_train_step(mipt_ctx* ctx, const float* d_queries, const float* d_targets, uint32_t n,
                                   float* out_loss, void* stream);
/* Gradients of the loss on n records, without an update (for gradient
 * checks): host_mlp_grad float[weights] in the [K][N] layout of each layer,
 * host_grid_grad float[levels * 2^log2_table * 2] (either may be NULL).
 * Synchronous; the accumulated gradients are cleared afterwards. */
int mipt_radiance_cache_gradients(mipt_ctx* ctx, const float* d_queries, const float* d_targets, uint32_t n,
                                  float* host_mlp_grad, float* host_grid_grad, float* out_loss);
/* Parameter readback / upload (host arrays; any may be NULL). MLP:
 * float master and bf16 copy, [K][N] per layer, 5 x 64 x 64 + 64 x 16.
 * Grid: float master and FP16 copy, [level][entry][feature]. */
int mipt_radiance_cache_get_params(mipt_ctx* ctx, float* mlp_master, uint16_t* mlp_bf16, float* grid_master,
                                   uint16_t* grid_f16);
int mipt_radiance_cache_set_params(mipt_ctx* ctx, const float* mlp_master, const float* grid_master);
int32_t mipt_radiance_cache_weight_count(void);
/* After a path pass with the cache: counts[0] queries answered by the cache,
 * counts[1] training records, counts[2] training targets dropped as non-finite
 * or negative; *train_loss the mean relative L2 loss of the last
 * RADIANCE_CACHE_TRAIN pass over its records (either may be NULL).
 * Synchronous. */
int mipt_radiance_cache_stats(mipt_ctx* ctx, uint32_t* counts, float* train_loss);

/* The path pass with render_settings.radiance_cache = 1 stops paths at
 * query_bounce and adds throughput x the cache's prediction; one path in
 * train_every continues train_extension more bounces and records a training
 * target. RADIANCE_CACHE_TRAIN then takes one optimiser step per batch_size
 * records of the last path pass. */
#define MIPT_PASS_RADIANCE_CACHE_TRAIN 4
/* A path pass that traces only training paths (one in warmup_train_every)
 * and leaves the framebuffer and its sample count alone: training frames
 * that cost a fraction of a rendered frame, for a cache that starts cold. */
#define MIPT_PASS_RADIANCE_CACHE_WARMUP 5
/* One warm-up frame: the passes of mipt_render_sample with
 * RADIANCE_CACHE_WARMUP in place of PATH_TRACE, then RADIANCE_CACHE_TRAIN.
 * Needs render_settings.radiance_cache = 1. */
int mipt_radiance_cache_warmup(mipt_ctx* ctx, uint32_t sample_number, void* stream);

/* Reservoir readback: which = 0 initial-candidates buffer, 1 / 2 the spatial
 * output buffers, 3 the reservoirs the path pass shades. Per pixel 8 words:
 * float point[3], float target_function, int32 M, int32 emissive_triangle,
 * float UCW, int32 flags (ReSTIRDIReservoir, DI/Reservoir.h:52-187, without
 * weight_sum, which no later pass reads). Synchronous. */
int mipt_read_restir_reservoirs(mipt_ctx* ctx, int32_t which, void* host_out);

/* Framebuffer: ColorRGB32F[W*H] SUM of samples (not the mean), pix = x + y*W,
 * y = 0 the bottom row (PathTracing.h:176-217). host_rgb: float[3*W*H] or NULL;
 * host_rgb64: double[3*W*H] or NULL (needs accumulate_fp64). Synchronous. */
int mipt_read_framebuffer(mipt_ctx* ctx, float* host_rgb, double* host_rgb64);

/* Device pointers of the framebuffer sums (valid until the next
 * mipt_set_render_settings). */
int mipt_framebuffer_device(mipt_ctx* ctx, float** d_rgb, double** d_rgb64);

/* G-buffer readback (first-hit primitive per pixel, -1 on a miss). */
int mipt_read_gbuffer_prims(mipt_ctx* ctx, int32_t* host_prims);

/* Samples dropped by the NaN/negative check (SanityCheck.h:78-106) since the
 * last mipt_set_render_settings. */
int mipt_rejected_samples(mipt_ctx* ctx, uint32_t* out_count);

/* ------------------------------------------------------------------ */
/* Procedural scenes (replace the ASSIMP SceneParser, which is out of   */
/* scope). config 1: Cornell box + 20 icospheres (102,428 triangles);  */
/* config 2: fBm terrain + 400 spheres + 4096 emitters (1,040,384);     */
/* config 5: 8 fBm grids + 3000 instances of 3 meshes (flattened) +     */
/* 16,384 emitters, every Principled layer (about 10M triangles).       */
/* Seeded PCG32 (state 0x2026120200000000 + config, stream = config).  */
/* ------------------------------------------------------------------ */
typedef struct mipt_generated_scene {
    mipt_scene_desc desc;   /* pointers owned by the generated scene */
    mipt_camera camera;     /* camera for (width, height) */
    int32_t width, height;
} mipt_generated_scene;

int mipt_scene_generate(int32_t config_id, int32_t width, int32_t height, mipt_generated_scene** out);
void mipt_scene_free(mipt_generated_scene* scene);

/* glTF 2.0 ingest: SceneParser::parse_scene_file (src/Scene/SceneParser.cpp:
 * 19-226) for a .gltf (external or base64 data: buffers) or .glb file, which
 * the reference reads through ASSIMP with aiProcess_PreTransformVertices |
 * aiProcess_Triangulate. Vertices and normals are pre-transformed to world
 * space (double precision, rounded once); triangles are grouped material by
 * material, within a material in depth-first node order; materials follow
 * SceneParser::read_material_properties (:392-436) over the importer's keys
 * (base colour, emission x KHR_materials_emissive_strength, metallic,
 * roughness, KHR_materials_anisotropy / sheen / specular / clearcoat / ior /
 * transmission / volume), plus the default glTF material, last, for
 * primitives without one; every triangle of a material with a non-zero
 * emission colour is an emitter. The camera is the first one, at its node's
 * origin looking down the node's -Z with +Y up, with its yfov at
 * width / height (without a camera: the origin looking down -Z, 40 degrees
 * horizontal). Textures, texture coordinates, skins, morph targets,
 * animations, point and line primitives, sparse and quantised accessors are
 * not read. Freed with mipt_scene_free; errors in mipt_last_error(NULL). */
int mipt_scene_load_gltf(const char* path, int32_t width, int32_t height, mipt_generated_scene** out);

/* Host xorshift32 seeded with 42, one draw per pass: the reference's
 * per-pass random_number (src/Renderer/GPURenderer.cpp:750-751). */
uint32_t mipt_host_random_number(uint32_t pass_index);

#ifdef __cplusplus
}
#endif
#endif
