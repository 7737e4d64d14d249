"""ctypes binding of libmipt.so (include/mipt.h) for tests and bench.py.

This module is plumbing: it loads the C-ABI library, mirrors its structs and
moves arrays between numpy and device memory (torch tensors on the HIP
device). All compute happens in libmipt's HIP kernels; if the library is not
built, loading fails loudly -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PACKAGE_DIR = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PACKAGE_DIR, "libmipt.so")

MIPT_OK = 0
MIPT_BVH_LBVH = 0
MIPT_SHADE_ORDER_LIVE = 0
MIPT_SHADE_ORDER_MATERIAL = 1
MIPT_SHADE_ORDER_AUTO = 2
MIPT_BVH_SAH_BINNED = 1

RAY_DTYPE = np.dtype([("o", "<f4", (3,)), ("last_prim", "<i4"), ("d", "<f4", (3,)), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("prim", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16

_F3 = C.c_float * 3


class Camera(C.Structure):
    """mipt_camera = HIPRTCamera (src/HostDeviceCommon/HIPRTCamera.h:16-27)."""

    _fields_ = [
        ("inverse_view", (C.c_float * 4) * 4),
        ("inverse_projection", (C.c_float * 4) * 4),
        ("view_projection", (C.c_float * 4) * 4),
        ("position", _F3),
        ("vertical_fov", C.c_float),
        ("sensor_width", C.c_int32),
        ("sensor_height", C.c_int32),
        ("do_jittering", C.c_int32),
    ]


_MATERIAL_FIELDS = [
    ("emission", _F3), ("emission_strength", C.c_float), ("global_emissive_factor", C.c_float),
    ("emissive_texture_used", C.c_int32), ("base_color", _F3), ("roughness", C.c_float),
    ("oren_nayar_sigma", C.c_float), ("metallic", C.c_float), ("metallic_F90_falloff_exponent", C.c_float),
    ("metallic_F82", _F3), ("metallic_F90", _F3), ("anisotropy", C.c_float), ("anisotropy_rotation", C.c_float),
    ("second_roughness_weight", C.c_float), ("second_roughness", C.c_float),
    ("do_metallic_energy_compensation", C.c_int32), ("specular", C.c_float), ("specular_tint", C.c_float),
    ("specular_color", _F3), ("specular_darkening", C.c_float), ("do_specular_energy_compensation", C.c_int32),
    ("coat", C.c_float), ("coat_medium_absorption", _F3), ("coat_medium_thickness", C.c_float),
    ("coat_roughness", C.c_float), ("coat_roughening", C.c_float), ("coat_darkening", C.c_float),
    ("coat_anisotropy", C.c_float), ("coat_anisotropy_rotation", C.c_float), ("coat_ior", C.c_float),
    ("do_coat_energy_compensation", C.c_int32), ("sheen", C.c_float), ("sheen_roughness", C.c_float),
    ("sheen_color", _F3), ("ior", C.c_float), ("specular_transmission", C.c_float),
    ("diffuse_transmission", C.c_float), ("absorption_at_distance", C.c_float), ("absorption_color", _F3),
    ("dispersion_scale", C.c_float), ("dispersion_abbe_number", C.c_float), ("thin_walled", C.c_int32),
    ("do_glass_energy_compensation", C.c_int32), ("thin_film", C.c_float), ("thin_film_ior", C.c_float),
    ("thin_film_thickness", C.c_float), ("thin_film_kappa_3", C.c_float),
    ("thin_film_hue_shift_degrees", C.c_float), ("thin_film_base_ior_override", C.c_float),
    ("thin_film_do_ior_override", C.c_int32), ("alpha_opacity", C.c_float), ("dielectric_priority", C.c_int32),
]


class Material(C.Structure):
    """mipt_material = CPUMaterial (src/HostDeviceCommon/Material/MaterialCPU.h:213-330)."""

    _fields_ = _MATERIAL_FIELDS


class SceneDesc(C.Structure):
    """mipt_scene_desc = struct Scene (src/Scene/SceneParser.h:109-170)."""

    _fields_ = [
        ("triangles_vertex_indices", C.POINTER(C.c_int32)),
        ("triangle_count", C.c_int32),
        ("vertices_positions", C.POINTER(C.c_float)),
        ("vertex_count", C.c_int32),
        ("has_vertex_normals", C.POINTER(C.c_uint8)),
        ("vertex_normals", C.POINTER(C.c_float)),
        ("material_indices", C.POINTER(C.c_int32)),
        ("materials", C.POINTER(Material)),
        ("material_count", C.c_int32),
        ("emissive_triangles_primitive_indices", C.POINTER(C.c_int32)),
        ("emissive_triangle_count", C.c_int32),
    ]


class View(C.Structure):
    """mipt_view: one camera view of a batch."""

    _fields_ = [("camera", Camera), ("width", C.c_int32), ("height", C.c_int32)]


class TraceTuning(C.Structure):
    """mipt_trace_tuning: the autotuner's choice."""

    _fields_ = [("block_size", C.c_int32), ("lds_stack", C.c_int32), ("queue_tile", C.c_int32),
                ("configurations", C.c_int32), ("best_ms", C.c_double), ("default_ms", C.c_double)]


class BvhInfo(C.Structure):
    _fields_ = [
        ("builder", C.c_int32), ("triangle_count", C.c_int32), ("node_count", C.c_int32),
        ("leaf_count", C.c_int32), ("max_depth", C.c_int32), ("max_leaf_size", C.c_int32),
        ("node_bytes", C.c_int64), ("triangle_bytes", C.c_int64), ("build_ms", C.c_double),
        ("sah_cost", C.c_double),
    ]


class RenderSettings(C.Structure):
    """mipt_render_settings (subset of HIPRTRenderSettings, RenderSettings.h:26-278)."""

    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("nb_bounces", C.c_int32),
        ("do_russian_roulette", C.c_int32), ("russian_roulette_min_depth", C.c_int32),
        ("russian_roulette_throughput_clamp", C.c_float), ("freeze_random", C.c_int32),
        ("enable_direct", C.c_int32), ("microfacet_regularization_tau_0", C.c_float),
        ("microfacet_regularization_min_roughness", C.c_float), ("do_microfacet_regularization", C.c_int32),
        ("accumulate_fp64", C.c_int32), ("direct_light_sampling_strategy", C.c_int32),
        ("direct_light_sampling_base_strategy", C.c_int32), ("radiance_cache", C.c_int32),
        ("do_normal_mapping", C.c_int32), ("accumulate", C.c_int32),
        ("ris_number_of_light_candidates", C.c_int32), ("ris_number_of_bsdf_candidates", C.c_int32),
    ]


class RestirDISettings(C.Structure):
    """mipt_restir_di_settings (ReSTIRDISettings.h:17-163)."""

    _fields_ = [
        ("number_of_initial_light_candidates", C.c_int32), ("number_of_initial_bsdf_candidates", C.c_int32),
        ("do_temporal_reuse_pass", C.c_int32), ("do_spatial_reuse_pass", C.c_int32),
        ("spatial_number_of_passes", C.c_int32), ("spatial_reuse_radius", C.c_int32),
        ("spatial_reuse_neighbor_count", C.c_int32), ("m_cap", C.c_int32), ("use_confidence_weights", C.c_int32),
        ("use_normal_similarity_heuristic", C.c_int32), ("normal_similarity_angle_precomp", C.c_float),
        ("use_plane_distance_heuristic", C.c_int32), ("plane_distance_threshold", C.c_float),
        ("reject_using_geometric_normals", C.c_int32), ("do_final_shading_visibility", C.c_int32),
        ("guard_backfacing_bsdf_candidates", C.c_int32), ("do_fused_spatiotemporal", C.c_int32),
        ("do_light_presampling", C.c_int32), ("light_presampling_number_of_subsets", C.c_int32),
        ("light_presampling_subset_size", C.c_int32), ("light_presampling_tile_size", C.c_int32),
        ("envmap_candidate_probability", C.c_float), ("use_permutation_sampling", C.c_int32),
        ("permutation_sampling_random_bits", C.c_int32), ("max_neighbor_search_count", C.c_int32),
        ("neighbor_search_radius", C.c_int32), ("later_bounces_sampling_strategy", C.c_int32),
    ]


class RadianceCacheSettings(C.Structure):
    """mipt_radiance_cache_settings (the learned radiance cache; no reference counterpart)."""

    _fields_ = [
        ("grid_levels", C.c_int32), ("grid_log2_table", C.c_int32), ("grid_base_resolution", C.c_int32),
        ("grid_per_level_scale", C.c_float), ("batch_size", C.c_int32), ("learning_rate", C.c_float),
        ("adam_beta1", C.c_float), ("adam_beta2", C.c_float), ("adam_eps", C.c_float), ("seed", C.c_uint32),
        ("query_bounce", C.c_int32), ("train_every", C.c_int32), ("train_extension", C.c_int32),
        ("warmup_train_every", C.c_int32),
    ]


MIPT_PASS_CAMERA_RAYS = 0
MIPT_PASS_PATH_TRACE = 1
MIPT_PASS_RESTIR_DI_INITIAL_CANDIDATES = 2
MIPT_PASS_RESTIR_DI_TEMPORAL_REUSE = 3
MIPT_PASS_RADIANCE_CACHE_TRAIN = 4
MIPT_PASS_RADIANCE_CACHE_WARMUP = 5
MIPT_PASS_RESTIR_DI_LIGHTS_PRESAMPLING = 6
MIPT_PASS_RESTIR_DI_SPATIOTEMPORAL_REUSE = 7


def MIPT_PASS_RESTIR_DI_SPATIAL_REUSE(i: int) -> int:
    return 16 + i


MIPT_LSS_MIS_LIGHT_BSDF = 3
MIPT_LSS_RIS_BSDF_AND_LIGHT = 4
MIPT_LSS_RESTIR_DI = 5
MIPT_RESTIR_DI_LATER_BOUNCES_MIS_LIGHT_BSDF = 2
MIPT_RESTIR_DI_LATER_BOUNCES_RIS_BSDF_AND_LIGHT = 3
MIPT_LSS_BASE_UNIFORM = 0
MIPT_LSS_BASE_POWER = 1

MIPT_AMBIENT_NONE = 0
MIPT_AMBIENT_UNIFORM = 1
MIPT_AMBIENT_ENVMAP = 2


class WorldSettings(C.Structure):
    """mipt_world_settings (WorldSettings.h:21-53)."""

    _fields_ = [
        ("ambient_light_type", C.c_int32), ("uniform_light_color", C.c_float * 3), ("envmap_intensity", C.c_float),
        ("envmap_scale_background_intensity", C.c_int32), ("envmap_to_world", (C.c_float * 3) * 3),
        ("world_to_envmap", (C.c_float * 3) * 3),
    ]


# Energy-compensation LUT baker (GPUBaker.cpp, src/Device/kernels/Baking/)
MIPT_LUT_GGX_CONDUCTOR = 0
MIPT_LUT_GLOSSY_DIELECTRIC = 1
MIPT_LUT_GGX_GLASS = 2
MIPT_LUT_GGX_GLASS_INVERSE = 3
MIPT_LUT_GGX_THIN_GLASS = 4
MIPT_LUT_GGX_FRESNEL = 5
MIPT_GGX_HEIGHT_CORRELATED = 0
MIPT_GGX_HEIGHT_UNCORRELATED = 1


class BakeSettings(C.Structure):
    """mipt_bake_settings (the *DirectionalAlbedoSettings structs)."""

    _fields_ = [
        ("table", C.c_int32), ("size_cos_theta", C.c_int32), ("size_roughness", C.c_int32),
        ("size_ior", C.c_int32), ("integration_sample_count", C.c_int32), ("masking_shadowing", C.c_int32),
        ("lanes_per_texel", C.c_int32),
    ]

    @property
    def shape(self):
        """(size_ior, size_roughness, size_cos_theta): the table's array shape."""
        return (self.size_ior, self.size_roughness, self.size_cos_theta)


class BakeInfo(C.Structure):
    _fields_ = [
        ("iterations_per_launch", C.c_int32), ("launches", C.c_int32), ("normalizer", C.c_int32),
        ("lanes_per_texel", C.c_int32), ("kernel_launches", C.c_int32), ("ms", C.c_float), ("samples", C.c_double),
    ]


# ReSTIR DI reservoir record as read back by mipt_read_restir_reservoirs.
RESERVOIR_DTYPE = np.dtype([("point", np.float32, 3), ("target", np.float32), ("M", np.int32),
                            ("tri", np.int32), ("ucw", np.float32), ("flags", np.int32)])
RESTIR_FLAG_UNOCCLUDED = 1 << 7


class GeneratedScene(C.Structure):
    _fields_ = [("desc", SceneDesc), ("camera", Camera), ("width", C.c_int32), ("height", C.c_int32)]


# Every extern "C" symbol of include/mipt.h with its ctypes signature.
_SIGNATURES = {
    "mipt_version": (C.c_char_p, []),
    "mipt_material_default": (None, [C.POINTER(Material)]),
    "mipt_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "mipt_destroy": (None, [C.c_void_p]),
    "mipt_last_error": (C.c_char_p, [C.c_void_p]),
    "mipt_set_scene": (C.c_int, [C.c_void_p, C.POINTER(SceneDesc)]),
    "mipt_build_bvh": (C.c_int, [C.c_void_p, C.c_int32]),
    "mipt_get_bvh_info": (C.c_int, [C.c_void_p, C.POINTER(BvhInfo)]),
    "mipt_read_bvh": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "mipt_trace_closest": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "mipt_trace_any": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "mipt_trace_closest_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "mipt_trace_closest_counted": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                             C.c_void_p]),
    "mipt_set_trace_config": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "mipt_set_shading_order": (C.c_int, [C.c_void_p, C.c_int32]),
    "mipt_set_views": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    "mipt_views_generate": (C.c_int, [C.c_int32, C.c_void_p]),
    "mipt_set_graph_capture": (C.c_int, [C.c_void_p, C.c_int32]),
    "mipt_graph_replays": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "mipt_set_queue_tile": (C.c_int, [C.c_void_p, C.c_int32]),
    "mipt_autotune_trace": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int32, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "mipt_update_materials": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    "mipt_update_material": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "mipt_camera_rays": (C.c_int, [C.c_void_p, C.POINTER(Camera), C.c_int32, C.c_int32, C.c_uint32, C.c_uint32,
                                   C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mipt_bounce_rays": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                   C.c_void_p]),
    "mipt_scene_generate": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.POINTER(GeneratedScene))]),
    "mipt_scene_free": (None, [C.POINTER(GeneratedScene)]),
    "mipt_scene_load_gltf": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.POINTER(C.POINTER(GeneratedScene))]),
    "mipt_host_random_number": (C.c_uint32, [C.c_uint32]),
    "mipt_render_settings_default": (None, [C.POINTER(RenderSettings)]),
    "mipt_set_render_settings": (C.c_int, [C.c_void_p, C.POINTER(RenderSettings)]),
    "mipt_set_camera": (C.c_int, [C.c_void_p, C.POINTER(Camera)]),
    "mipt_launch_pass": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32, C.c_void_p]),
    "mipt_render_sample": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "mipt_radiance_cache_warmup": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "mipt_read_framebuffer": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "mipt_framebuffer_device": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "mipt_read_gbuffer_prims": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mipt_rejected_samples": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "mipt_restir_di_settings_default": (None, [C.POINTER(RestirDISettings)]),
    "mipt_set_restir_di_settings": (C.c_int, [C.c_void_p, C.POINTER(RestirDISettings)]),
    "mipt_draws_per_sample": (C.c_int32, [C.c_void_p]),
    "mipt_get_power_alias_table": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    "mipt_world_settings_default": (None, [C.POINTER(WorldSettings)]),
    "mipt_set_world_settings": (C.c_int, [C.c_void_p, C.POINTER(WorldSettings)]),
    "mipt_get_world_settings": (C.c_int, [C.c_void_p, C.POINTER(WorldSettings)]),
    "mipt_set_envmap": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "mipt_get_envmap_tables": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_float), C.c_void_p]),
    "mipt_envmap_generate": (C.c_int, [C.c_int32, C.c_int32, C.c_void_p]),
    "mipt_bake_settings_default": (C.c_int, [C.c_int32, C.POINTER(BakeSettings)]),
    "mipt_bake_lut": (C.c_int, [C.c_void_p, C.POINTER(BakeSettings), C.c_void_p, C.POINTER(BakeInfo)]),
    "mipt_set_lut": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "mipt_get_shipped_lut": (C.c_int, [C.c_int32, C.c_int32, C.c_void_p, C.c_int32]),
    "mipt_compute_alias_table": (C.c_int, [C.c_void_p, C.c_int32, C.c_float, C.c_void_p, C.c_void_p]),
    "mipt_read_restir_reservoirs": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "mipt_radiance_cache_settings_default": (None, [C.POINTER(RadianceCacheSettings)]),
    "mipt_set_radiance_cache": (C.c_int, [C.c_void_p, C.POINTER(RadianceCacheSettings)]),
    "mipt_radiance_cache_infer": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "mipt_radiance_cache_encode": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "mipt_radiance_cache_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                                 C.POINTER(C.c_float), C.c_void_p]),
    "mipt_radiance_cache_gradients": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                                C.c_void_p, C.POINTER(C.c_float)]),
    "mipt_radiance_cache_get_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mipt_radiance_cache_set_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "mipt_radiance_cache_weight_count": (C.c_int32, []),
    "mipt_radiance_cache_stats": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libmipt.so; raises if it is not built (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libmipt.so not built at {path}: run __graft_entry__.build()")
    # torch ships its own libamdhip64.so.7. Load it first so that libmipt's
    # NEEDED entry binds to that same runtime: two HIP runtimes in one process
    # leave the second one without a usable device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return list(_SIGNATURES)


def host_random_number(pass_index: int) -> int:
    return int(load_library().mipt_host_random_number(pass_index))


def default_render_settings(**overrides) -> RenderSettings:
    rs = RenderSettings()
    load_library().mipt_render_settings_default(C.byref(rs))
    for k, v in overrides.items():
        setattr(rs, k, v)
    return rs


def default_restir_di_settings(**overrides) -> RestirDISettings:
    rs = RestirDISettings()
    load_library().mipt_restir_di_settings_default(C.byref(rs))
    for k, v in overrides.items():
        setattr(rs, k, v)
    return rs


def default_world_settings(**overrides) -> WorldSettings:
    w = WorldSettings()
    load_library().mipt_world_settings_default(C.byref(w))
    for k, v in overrides.items():
        if k == "uniform_light_color":
            w.uniform_light_color[:] = [float(x) for x in v]
        elif k in ("envmap_to_world", "world_to_envmap"):
            m = np.asarray(v, np.float32).reshape(3, 3)
            for c in range(3):
                getattr(w, k)[c][:] = [float(x) for x in m[c]]
        else:
            setattr(w, k, v)
    return w


def default_bake_settings(table: int, **overrides) -> BakeSettings:
    """The reference's bake settings of a table at the renderer's size."""
    b = BakeSettings()
    rc = load_library().mipt_bake_settings_default(table, C.byref(b))
    if rc != MIPT_OK:
        raise MiptError(rc, "mipt_bake_settings_default")
    for k, v in overrides.items():
        setattr(b, k, v)
    return b


def shipped_lut(table: int, masking_shadowing: int = MIPT_GGX_HEIGHT_UNCORRELATED) -> np.ndarray:
    """A table the library ships (the reference's baked files), [z][roughness
    flipped][cos_theta]."""
    b = default_bake_settings(table)
    out = np.zeros(b.shape, np.float32)
    rc = load_library().mipt_get_shipped_lut(table, masking_shadowing, out.ctypes.data_as(C.c_void_p), out.size)
    if rc != MIPT_OK:
        raise MiptError(rc, "mipt_get_shipped_lut")
    return out


def generate_envmap(width: int, height: int) -> np.ndarray:
    """The procedural sun + sky envmap (mipt_envmap_generate), height x width x 3."""
    out = np.zeros((height, width, 3), np.float32)
    rc = load_library().mipt_envmap_generate(width, height, out.ctypes.data_as(C.c_void_p))
    if rc != MIPT_OK:
        raise MiptError(rc, "mipt_envmap_generate")
    return out


def compute_alias_table(weights, total_sum=None):
    """Vose alias table of `weights` (FP64 normalisation), host code of libmipt."""
    w = np.ascontiguousarray(weights, np.float32)
    if total_sum is None:
        total_sum = np.float32(0.0)
        for x in w:  # float accumulation in order, like the reference's power sum
            total_sum = np.float32(total_sum + x)
    probas = np.zeros(w.size, np.float32)
    alias = np.zeros(w.size, np.int32)
    rc = load_library().mipt_compute_alias_table(w.ctypes.data_as(C.c_void_p), w.size, float(total_sum),
                                                 probas.ctypes.data_as(C.c_void_p), alias.ctypes.data_as(C.c_void_p))
    if rc != MIPT_OK:
        raise MiptError(rc, "mipt_compute_alias_table")
    return probas, alias


def default_radiance_cache_settings(**overrides) -> RadianceCacheSettings:
    s = RadianceCacheSettings()
    load_library().mipt_radiance_cache_settings_default(C.byref(s))
    for k, v in overrides.items():
        setattr(s, k, v)
    return s


def default_material() -> Material:
    m = Material()
    load_library().mipt_material_default(C.byref(m))
    return m


class MiptError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mipt error {code}: {msg}")
        self.code = code


# ---------------------------------------------------------------------------
# Procedural scenes
# ---------------------------------------------------------------------------
class Scene:
    """Host arrays of a scene (numpy), in the layout of struct Scene."""

    def __init__(self, vertices, indices, material_indices=None, materials=None, emissive=None,
                 camera=None, width=0, height=0, vertex_normals=None, has_vertex_normals=None):
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        # smooth shading: float32 (V, 3) normals and a uint8 (V,) flag per vertex
        self.vertex_normals = None if vertex_normals is None else \
            np.ascontiguousarray(vertex_normals, dtype=np.float32).reshape(-1, 3)
        self.has_vertex_normals = None
        if self.vertex_normals is not None:
            self.has_vertex_normals = np.ones(self.vertices.shape[0], np.uint8) if has_vertex_normals is None else \
                np.ascontiguousarray(has_vertex_normals, dtype=np.uint8)
        self.indices = np.ascontiguousarray(indices, dtype=np.int32).reshape(-1, 3)
        self.material_indices = None if material_indices is None else np.ascontiguousarray(material_indices,
                                                                                              dtype=np.int32)
        self.materials = materials
        self.emissive = np.zeros(0, np.int32) if emissive is None else np.ascontiguousarray(emissive, np.int32)
        self.camera = camera
        self.width = width
        self.height = height

    @property
    def triangle_count(self):
        return int(self.indices.shape[0])

    def desc(self) -> SceneDesc:
        d = SceneDesc()
        d.triangles_vertex_indices = self.indices.ctypes.data_as(C.POINTER(C.c_int32))
        d.triangle_count = self.indices.shape[0]
        d.vertices_positions = self.vertices.ctypes.data_as(C.POINTER(C.c_float))
        d.vertex_count = self.vertices.shape[0]
        d.has_vertex_normals = None
        d.vertex_normals = None
        if self.vertex_normals is not None:
            d.has_vertex_normals = self.has_vertex_normals.ctypes.data_as(C.POINTER(C.c_uint8))
            d.vertex_normals = self.vertex_normals.ctypes.data_as(C.POINTER(C.c_float))
        if self.material_indices is not None and self.materials is not None:
            d.material_indices = self.material_indices.ctypes.data_as(C.POINTER(C.c_int32))
            d.materials = C.cast(self.materials, C.POINTER(Material))
            d.material_count = len(self.materials)
        d.emissive_triangles_primitive_indices = self.emissive.ctypes.data_as(C.POINTER(C.c_int32))
        d.emissive_triangle_count = self.emissive.shape[0]
        return d


def generate_views(count: int):
    """The seeded view batch of configuration 3 (mipt_views_generate)."""
    views = (View * count)()
    rc = load_library().mipt_views_generate(count, C.cast(views, C.c_void_p))
    if rc != MIPT_OK:
        raise MiptError(rc, "mipt_views_generate")
    return views


def generate_scene(config_id: int, width: int, height: int) -> Scene:
    """Procedural scene of configuration 1, 2 or 5, generated by libmipt."""
    lib = load_library()
    out = C.POINTER(GeneratedScene)()
    rc = lib.mipt_scene_generate(config_id, width, height, C.byref(out))
    if rc != MIPT_OK:
        raise MiptError(rc, f"mipt_scene_generate({config_id})")
    return _take_scene(lib, out)


def load_gltf(path: str, width: int, height: int) -> Scene:
    """A glTF 2.0 scene (.gltf / .glb) read by libmipt (mipt_scene_load_gltf),
    with its vertex normals and camera."""
    lib = load_library()
    out = C.POINTER(GeneratedScene)()
    rc = lib.mipt_scene_load_gltf(os.fsencode(path), width, height, C.byref(out))
    if rc != MIPT_OK:
        raise MiptError(rc, "mipt_scene_load_gltf: " + lib.mipt_last_error(None).decode())
    return _take_scene(lib, out)


def _take_scene(lib, out) -> Scene:
    """Copies a mipt_generated_scene into a Scene and frees it."""
    try:
        g = out.contents
        d = g.desc
        T, V, M, E = d.triangle_count, d.vertex_count, d.material_count, d.emissive_triangle_count
        idx = np.ctypeslib.as_array(d.triangles_vertex_indices, shape=(3 * T,)).copy()
        verts = np.ctypeslib.as_array(d.vertices_positions, shape=(3 * V,)).copy()
        mat_idx = np.ctypeslib.as_array(d.material_indices, shape=(T,)).copy()
        mats = (Material * M)()
        C.memmove(mats, d.materials, C.sizeof(Material) * M)
        emissive = np.ctypeslib.as_array(d.emissive_triangles_primitive_indices, shape=(E,)).copy() if E else None
        normals = has_normals = None
        if d.vertex_normals and d.has_vertex_normals:
            normals = np.ctypeslib.as_array(d.vertex_normals, shape=(3 * V,)).copy()
            has_normals = np.ctypeslib.as_array(d.has_vertex_normals, shape=(V,)).copy()
        cam = Camera()
        C.memmove(C.byref(cam), C.byref(g.camera), C.sizeof(Camera))
        return Scene(verts, idx, mat_idx, mats, emissive, cam, g.width, g.height, vertex_normals=normals,
                     has_vertex_normals=has_normals)
    finally:
        lib.mipt_scene_free(out)


# ---------------------------------------------------------------------------
# Device memory helpers (torch is only used as the HIP allocator)
# ---------------------------------------------------------------------------
def _torch():
    import torch

    return torch


def to_device(arr: np.ndarray):
    torch = _torch()
    raw = np.ascontiguousarray(arr).view(np.uint8).reshape(-1)
    return torch.from_numpy(raw.copy()).to("cuda")


def empty_device(nbytes: int):
    torch = _torch()
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device="cuda")


def from_device(t, dtype) -> np.ndarray:
    return t.cpu().numpy().view(dtype)


def ptr(t) -> int:
    return int(t.data_ptr())


class Context:
    """One mipt context on one HIP device."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.mipt_create(device, C.byref(h))
        if rc != MIPT_OK:
            msg = self.lib.mipt_last_error(None)
            raise MiptError(rc, f"mipt_create failed: {msg.decode() if msg else ''}")
        self.h = h

    def close(self):
        if self.h:
            self.lib.mipt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != MIPT_OK:
            msg = self.lib.mipt_last_error(self.h)
            raise MiptError(rc, f"{what}: {msg.decode() if msg else ''}")

    def set_scene(self, scene: Scene):
        d = scene.desc()
        self._check(self.lib.mipt_set_scene(self.h, C.byref(d)), "mipt_set_scene")

    def build_bvh(self, builder: int = MIPT_BVH_SAH_BINNED):
        """Builds the scene BVH; by default with the top-down binned-SAH
        builder (the reference builds with HIPRT's high-quality flag,
        HIPRTScene.h:87-150), MIPT_BVH_LBVH for the faster LBVH build."""
        self._check(self.lib.mipt_build_bvh(self.h, builder), "mipt_build_bvh")

    def update_materials(self, materials):
        """GPURenderer::update_all_materials: a ctypes array of Material."""
        self._check(self.lib.mipt_update_materials(self.h, C.cast(materials, C.c_void_p), len(materials)),
                    "mipt_update_materials")

    def update_material(self, index: int, material):
        """GPURenderer::update_one_material."""
        self._check(self.lib.mipt_update_material(self.h, index, C.byref(material)), "mipt_update_material")

    def set_views(self, views):
        """mipt_set_views: a ctypes array of View (or an empty list for one view)."""
        if len(views) == 0:
            self._check(self.lib.mipt_set_views(self.h, None, 0), "mipt_set_views")
            self._views_total = 0
        else:
            self._check(self.lib.mipt_set_views(self.h, C.cast(views, C.c_void_p), len(views)), "mipt_set_views")
            self._views_total = sum(v.width * v.height for v in views)
            if getattr(self, "settings", None) is None:
                self.settings = RenderSettings()
                self.lib.mipt_render_settings_default(C.byref(self.settings))
            self.settings.width, self.settings.height = self._views_total, 1

    def set_graph_capture(self, enable: bool):
        self._check(self.lib.mipt_set_graph_capture(self.h, int(enable)), "mipt_set_graph_capture")

    def graph_replays(self) -> int:
        out = C.c_uint32()
        self._check(self.lib.mipt_graph_replays(self.h, C.byref(out)), "mipt_graph_replays")
        return int(out.value)

    def set_queue_tile(self, tile: int):
        self._check(self.lib.mipt_set_queue_tile(self.h, tile), "mipt_set_queue_tile")

    def autotune_trace(self, d_rays, n: int, repeats: int = 3, stream=None):
        """(TraceTuning, table[3 blocks][4 stacks][3 tiles] of mean ms)."""
        out = TraceTuning()
        table = np.zeros((3, 4, 3), np.float64)
        self._check(self.lib.mipt_autotune_trace(self.h, ptr(d_rays), n, repeats, stream, C.byref(out),
                                                 table.ctypes.data_as(C.c_void_p)), "mipt_autotune_trace")
        return out, table

    def set_shading_order(self, order: int):
        """MIPT_SHADE_ORDER_LIVE / _MATERIAL / _AUTO (the default)."""
        self._check(self.lib.mipt_set_shading_order(self.h, order), "mipt_set_shading_order")

    def bvh_info(self) -> BvhInfo:
        info = BvhInfo()
        self._check(self.lib.mipt_get_bvh_info(self.h, C.byref(info)), "mipt_get_bvh_info")
        return info

    def read_bvh(self):
        """(nodes, triangles): the BVH's 64-byte nodes as (n, 16) float32
        rows (boxes; the child refs are the int32 view of columns 12-13) and
        its 48-byte triangle records as (T, 12) float32 rows."""
        info = self.bvh_info()
        nodes = np.zeros((max(info.node_count, 0), 16), np.float32)
        tris = np.zeros((max(info.triangle_count, 0), 12), np.float32)
        self._check(self.lib.mipt_read_bvh(self.h, nodes.ctypes.data_as(C.c_void_p),
                                           tris.ctypes.data_as(C.c_void_p)), "mipt_read_bvh")
        return nodes, tris

    def set_trace_config(self, block_size: int = 0, lds_stack: int = 0):
        self._check(self.lib.mipt_set_trace_config(self.h, block_size, lds_stack), "mipt_set_trace_config")

    def trace_closest(self, d_rays, d_hits, n: int, stream=None):
        self._check(self.lib.mipt_trace_closest(self.h, ptr(d_rays), ptr(d_hits), n, stream),
                    "mipt_trace_closest")

    def trace_closest_noalloc(self, d_rays, d_hits, n: int, stream=None):
        """mipt_trace_closest through the library's static scratch: no device
        allocation (calls overlapping on one GPU race)."""
        self._check(self.lib.mipt_trace_closest_noalloc(self.h, ptr(d_rays), ptr(d_hits), n, stream),
                    "mipt_trace_closest_noalloc")

    def trace_closest_counted(self, d_rays, d_hits, d_counts, n: int, stream=None):
        self._check(self.lib.mipt_trace_closest_counted(self.h, ptr(d_rays), ptr(d_hits), ptr(d_counts), n, stream),
                    "mipt_trace_closest_counted")

    def trace_any(self, d_rays, d_occ, n: int, stream=None):
        self._check(self.lib.mipt_trace_any(self.h, ptr(d_rays), ptr(d_occ), n, stream), "mipt_trace_any")

    def camera_rays(self, cam: Camera, w: int, h: int, sample_number: int, random_number: int, freeze: int,
                    d_rays, d_rng, stream=None):
        self._check(self.lib.mipt_camera_rays(self.h, C.byref(cam), w, h, sample_number, random_number, freeze,
                                              ptr(d_rays), ptr(d_rng) if d_rng is not None else None, stream),
                    "mipt_camera_rays")

    def bounce_rays(self, d_primary, d_hits, d_rng, d_out, n: int, stream=None):
        self._check(self.lib.mipt_bounce_rays(self.h, ptr(d_primary), ptr(d_hits), ptr(d_rng), ptr(d_out), n,
                                              stream), "mipt_bounce_rays")

    # Path-tracing passes
    def set_render_settings(self, settings: RenderSettings):
        self._check(self.lib.mipt_set_render_settings(self.h, C.byref(settings)), "mipt_set_render_settings")
        self.settings = RenderSettings()
        C.memmove(C.byref(self.settings), C.byref(settings), C.sizeof(RenderSettings))
        if getattr(self, "_views_total", 0):
            # a batch of views: every view's pixels back to back
            self.settings.width, self.settings.height = self._views_total, 1

    def set_camera(self, cam: Camera):
        self._check(self.lib.mipt_set_camera(self.h, C.byref(cam)), "mipt_set_camera")

    def launch_pass(self, pass_id: int, sample_number: int, random_number: int, stream=None):
        self._check(self.lib.mipt_launch_pass(self.h, pass_id, sample_number, random_number, stream),
                    "mipt_launch_pass")

    def render_sample(self, sample_number: int, stream=None):
        self._check(self.lib.mipt_render_sample(self.h, sample_number, stream), "mipt_render_sample")

    def cache_warmup(self, sample_number: int, stream=None):
        """One radiance-cache warm-up frame: training paths only, then an
        optimiser step; the framebuffer is left alone."""
        self._check(self.lib.mipt_radiance_cache_warmup(self.h, sample_number, stream), "mipt_radiance_cache_warmup")

    def read_framebuffer(self, fp64: bool = False):
        n = self.settings.width * self.settings.height
        fb = np.zeros(3 * n, np.float32)
        fb64 = np.zeros(3 * n, np.float64) if fp64 else None
        self._check(self.lib.mipt_read_framebuffer(self.h, fb.ctypes.data_as(C.c_void_p),
                                                   fb64.ctypes.data_as(C.c_void_p) if fp64 else None),
                    "mipt_read_framebuffer")
        return (fb, fb64) if fp64 else fb

    def read_gbuffer_prims(self):
        n = self.settings.width * self.settings.height
        out = np.zeros(n, np.int32)
        self._check(self.lib.mipt_read_gbuffer_prims(self.h, out.ctypes.data_as(C.c_void_p)),
                    "mipt_read_gbuffer_prims")
        return out

    def set_restir_di_settings(self, settings: RestirDISettings):
        self._check(self.lib.mipt_set_restir_di_settings(self.h, C.byref(settings)), "mipt_set_restir_di_settings")
        self.restir = settings

    def draws_per_sample(self) -> int:
        return int(self.lib.mipt_draws_per_sample(self.h))

    def power_alias_table(self, emissive_count: int):
        probas = np.zeros(emissive_count, np.float32)
        alias = np.zeros(emissive_count, np.int32)
        total = C.c_float()
        self._check(self.lib.mipt_get_power_alias_table(self.h, probas.ctypes.data_as(C.c_void_p),
                                                        alias.ctypes.data_as(C.c_void_p), C.byref(total)),
                    "mipt_get_power_alias_table")
        return probas, alias, total.value

    # Ambient light
    def set_world_settings(self, world: WorldSettings):
        self._check(self.lib.mipt_set_world_settings(self.h, C.byref(world)), "mipt_set_world_settings")

    def world_settings(self) -> WorldSettings:
        w = WorldSettings()
        self._check(self.lib.mipt_get_world_settings(self.h, C.byref(w)), "mipt_get_world_settings")
        return w

    def set_envmap(self, image):
        """`image`: height x width x channels float32 (3 or 4 channels), or None to clear."""
        if image is None:
            self._check(self.lib.mipt_set_envmap(self.h, None, 0, 0, 3), "mipt_set_envmap")
            return
        img = np.ascontiguousarray(image, np.float32)
        h, w, c = img.shape
        self._check(self.lib.mipt_set_envmap(self.h, img.ctypes.data_as(C.c_void_p), w, h, c), "mipt_set_envmap")

    def bake_lut(self, settings: BakeSettings):
        """(table [z][roughness flipped][cos_theta] float32, BakeInfo) of a
        bake on the GPU (mipt_bake_lut)."""
        n = settings.size_cos_theta * settings.size_roughness * settings.size_ior
        dev = empty_device(4 * n)
        info = BakeInfo()
        self._check(self.lib.mipt_bake_lut(self.h, C.byref(settings), ptr(dev), C.byref(info)), "mipt_bake_lut")
        return from_device(dev, np.float32)[:n].reshape(settings.shape).copy(), info

    def set_lut(self, table: int, values=None):
        """Installs a table for the passes from then on (None: the shipped one)."""
        if values is None:
            self._check(self.lib.mipt_set_lut(self.h, table, None, 0, 0, 0), "mipt_set_lut")
            return
        v = np.ascontiguousarray(values, np.float32)
        shape = v.shape if v.ndim == 3 else (1,) + v.shape
        self._check(self.lib.mipt_set_lut(self.h, table, v.ctypes.data_as(C.c_void_p), shape[2], shape[1], shape[0]),
                    "mipt_set_lut")

    def envmap_tables(self, count: int):
        """(alias probabilities, aliases, luminance sum, packed R9G9B9E5 texels)."""
        probas = np.zeros(count, np.float32)
        alias = np.zeros(count, np.int32)
        texels = np.zeros(count, np.uint32)
        total = C.c_float()
        self._check(self.lib.mipt_get_envmap_tables(self.h, probas.ctypes.data_as(C.c_void_p),
                                                    alias.ctypes.data_as(C.c_void_p), C.byref(total),
                                                    texels.ctypes.data_as(C.c_void_p)), "mipt_get_envmap_tables")
        return probas, alias, total.value, texels

    def read_reservoirs(self, which: int = 3) -> np.ndarray:
        n = self.settings.width * self.settings.height
        out = np.zeros(n, RESERVOIR_DTYPE)
        self._check(self.lib.mipt_read_restir_reservoirs(self.h, which, out.ctypes.data_as(C.c_void_p)),
                    "mipt_read_restir_reservoirs")
        return out

    # Learned radiance cache
    def set_radiance_cache(self, settings: RadianceCacheSettings):
        self.cache = settings
        self._check(self.lib.mipt_set_radiance_cache(self.h, C.byref(settings)), "mipt_set_radiance_cache")

    def cache_infer(self, d_queries, n: int, d_out, stream=None):
        self._check(self.lib.mipt_radiance_cache_infer(self.h, ptr(d_queries), n, ptr(d_out), stream),
                    "mipt_radiance_cache_infer")

    def cache_encode(self, d_queries, n: int, d_out, stream=None):
        self._check(self.lib.mipt_radiance_cache_encode(self.h, ptr(d_queries), n, ptr(d_out), stream),
                    "mipt_radiance_cache_encode")

    def cache_train_step(self, d_queries, d_targets, n: int, want_loss: bool = True, stream=None):
        loss = C.c_float()
        self._check(self.lib.mipt_radiance_cache_train_step(self.h, ptr(d_queries), ptr(d_targets), n,
                                                            C.byref(loss) if want_loss else None, stream),
                    "mipt_radiance_cache_train_step")
        return loss.value if want_loss else None

    def cache_gradients(self, d_queries, d_targets, n: int):
        """(MLP gradient [K][N] per layer, grid gradient, loss) of one batch, without an update."""
        wg = np.zeros(int(self.lib.mipt_radiance_cache_weight_count()), np.float32)
        gg = np.zeros(self.cache.grid_levels * (1 << self.cache.grid_log2_table) * 2, np.float32)
        loss = C.c_float()
        self._check(self.lib.mipt_radiance_cache_gradients(self.h, ptr(d_queries), ptr(d_targets), n,
                                                           wg.ctypes.data_as(C.c_void_p),
                                                           gg.ctypes.data_as(C.c_void_p), C.byref(loss)),
                    "mipt_radiance_cache_gradients")
        return wg, gg, loss.value

    def cache_params(self):
        """(MLP FP32 master, MLP bf16 bits, grid FP32 master, grid FP16 bits)."""
        nw = int(self.lib.mipt_radiance_cache_weight_count())
        ng = self.cache.grid_levels * (1 << self.cache.grid_log2_table) * 2
        wm, wb = np.zeros(nw, np.float32), np.zeros(nw, np.uint16)
        gm, gh = np.zeros(ng, np.float32), np.zeros(ng, np.uint16)
        self._check(self.lib.mipt_radiance_cache_get_params(self.h, *(a.ctypes.data_as(C.c_void_p)
                                                                      for a in (wm, wb, gm, gh))),
                    "mipt_radiance_cache_get_params")
        return wm, wb, gm, gh

    def set_cache_params(self, mlp_master=None, grid_master=None):
        def p(a):
            return None if a is None else np.ascontiguousarray(a, np.float32).ctypes.data_as(C.c_void_p)
        keep = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (mlp_master, grid_master)]
        self._check(self.lib.mipt_radiance_cache_set_params(self.h, p(keep[0]), p(keep[1])),
                    "mipt_radiance_cache_set_params")

    def cache_stats(self):
        """(queries, training records, dropped targets, train loss) of the last path and train passes."""
        counts = np.zeros(3, np.uint32)
        loss = C.c_float()
        self._check(self.lib.mipt_radiance_cache_stats(self.h, counts.ctypes.data_as(C.c_void_p), C.byref(loss)),
                    "mipt_radiance_cache_stats")
        return int(counts[0]), int(counts[1]), int(counts[2]), loss.value

    def rejected_samples(self) -> int:
        v = C.c_uint32()
        self._check(self.lib.mipt_rejected_samples(self.h, C.byref(v)), "mipt_rejected_samples")
        return int(v.value)

    # Convenience wrappers taking and returning host arrays (not for timing).
    def trace_closest_host(self, rays: np.ndarray) -> np.ndarray:
        n = rays.shape[0]
        d_rays = to_device(rays)
        d_hits = empty_device(16 * n)
        self.trace_closest(d_rays, d_hits, n)
        _torch().cuda.synchronize()
        return from_device(d_hits, HIT_DTYPE)[:n]

    def trace_any_host(self, rays: np.ndarray) -> np.ndarray:
        n = rays.shape[0]
        d_rays = to_device(rays)
        d_occ = empty_device(n)
        self.trace_any(d_rays, d_occ, n)
        _torch().cuda.synchronize()
        return from_device(d_occ, np.uint8)[:n]
