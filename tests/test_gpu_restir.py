"""GPU parity of the ReSTIR DI passes (light presampling, initial
candidates, temporal reuse, spatial reuse, fused spatiotemporal reuse, final
shading at bounce 0) and of power light sampling,
through the C ABI, against the oracle's per-pixel restatement.

Bar: first-hit primitive IDs bit-exact; reservoirs (sample point, light
triangle, target function, M, UCW, flags) bit-identical on >= 99.9 % of
pixels; radiance within 1e-4 relative on >= 99.9 % of pixels (the sin/cos/pow
sequences are shared, so both usually agree bit for bit; any remaining
difference is counted and printed).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4
ABS_TOL = 1e-7
MIN_PASS = 0.999


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def simplify_materials(scene):
    """Config-2 materials restricted to the lobes the passes implement
    (Lambertian diffuse + the two GGX metal lobes)."""
    for m in scene.materials:
        m.coat = 0.0
        m.sheen = 0.0
        m.specular_transmission = 0.0
        m.diffuse_transmission = 0.0
        m.thin_film = 0.0
        m.specular = 0.0
    return scene


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32).reshape(a.shape[0], -1)


def _compare_fb(g, c):
    g = g.reshape(-1, 3).astype(np.float64)
    c = c.reshape(-1, 3).astype(np.float64)
    ok = (np.abs(g - c) <= REL_TOL * np.maximum(np.abs(g), np.abs(c)) + ABS_TOL).all(axis=1)
    return ok.mean(), int((~ok).sum())


def render_both(mipt_mod, oracle_mod, scene, settings, restir, frames, check_every_frame=False):
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    ctx.set_restir_di_settings(restir)
    ctx.set_render_settings(settings)
    ctx.set_camera(scene.camera)
    R = oracle_mod.Renderer(scene, settings, restir)
    draws = ctx.draws_per_sample()
    stats = []
    for f in range(frames):
        ctx.render_sample(f)
        R.render_frame(f, [mipt_mod.host_random_number(f * draws + k) for k in range(draws)])
        if check_every_frame or f == frames - 1:
            assert np.array_equal(ctx.read_gbuffer_prims(), R.gb_prims), f"frame {f}: first-hit IDs differ"
            for which in (0, 1, 2, 3):
                if settings.direct_light_sampling_strategy != mipt_mod.MIPT_LSS_RESTIR_DI:
                    break
                g = ctx.read_reservoirs(which)
                c = R.read_reservoirs(which)
                same = (_bits(g) == _bits(c)).all(axis=1)
                stats.append((f, which, same.mean()))
                assert same.mean() >= MIN_PASS, (f, which, same.mean(), np.nonzero(~same)[0][:5])
    fb, fb64 = ctx.read_framebuffer(fp64=True)
    frac, diverged = _compare_fb(fb, R.fb)
    frac64, _ = _compare_fb(fb64, R.fb64)
    identical = int((fb.reshape(-1, 3) == R.fb.reshape(-1, 3)).all(axis=1).sum())
    rej = ctx.rejected_samples()
    print(f"\n{settings.width}x{settings.height} x{frames}: radiance {frac * 100:.4f} % within 1e-4 ({diverged} "
          f"divergent, {identical}/{fb.size // 3} bit-identical); reservoirs identical "
          f"{min([s[2] for s in stats]) * 100 if stats else 100:.4f} %; rejected gpu {rej} cpu {R.rejected}")
    assert frac >= MIN_PASS and frac64 >= MIN_PASS
    assert rej == R.rejected
    # stronger than the bar: bit-identical frames (shared sin/cos/pow sequences)
    assert identical == fb.size // 3
    ctx.close()
    return fb, R


def test_restir_config1(mipt_mod, oracle_mod, torch_mod):
    scene = mipt_mod.generate_scene(1, 96, 96)
    s = mipt_mod.default_render_settings(width=96, height=96, direct_light_sampling_strategy=5)
    rs = mipt_mod.default_restir_di_settings()
    fb, R = render_both(mipt_mod, oracle_mod, scene, s, rs, 6, check_every_frame=True)
    assert fb.mean() > 0.01
    res = R.read_reservoirs(3)
    hit = R.gb_prims >= 0
    assert (res["M"][hit] >= 1).all() and (res["M"] <= rs.m_cap).all()


VARIANTS = {
    "no_temporal": dict(do_temporal_reuse_pass=0),
    "no_spatial": dict(do_spatial_reuse_pass=0),
    "initial_only": dict(do_temporal_reuse_pass=0, do_spatial_reuse_pass=0),
    "two_spatial": dict(spatial_number_of_passes=2),
    "three_spatial_no_temporal": dict(spatial_number_of_passes=3, do_temporal_reuse_pass=0),
    "no_confidence": dict(use_confidence_weights=0),
    "many_candidates": dict(number_of_initial_light_candidates=16, number_of_initial_bsdf_candidates=2, m_cap=20),
    "no_heuristics": dict(use_normal_similarity_heuristic=0, use_plane_distance_heuristic=0,
                          reject_using_geometric_normals=0),
    "no_final_visibility": dict(do_final_shading_visibility=0),
    "one_neighbor_radius3": dict(spatial_reuse_neighbor_count=1, spatial_reuse_radius=3),
    "fused": dict(do_fused_spatiotemporal=1),
    "fused_two_spatial": dict(do_fused_spatiotemporal=1, spatial_number_of_passes=2),
    "fused_three_no_confidence": dict(do_fused_spatiotemporal=1, spatial_number_of_passes=3,
                                      use_confidence_weights=0),
    "fused_spatial_flag_off": dict(do_fused_spatiotemporal=1, do_spatial_reuse_pass=0, spatial_number_of_passes=2),
    "fused_no_heuristics": dict(do_fused_spatiotemporal=1, use_normal_similarity_heuristic=0,
                                use_plane_distance_heuristic=0),
    "presampling": dict(do_light_presampling=1, light_presampling_number_of_subsets=16,
                        light_presampling_subset_size=64),
    "presampling_one_subset": dict(do_light_presampling=1, light_presampling_number_of_subsets=1,
                                   light_presampling_subset_size=7, light_presampling_tile_size=3),
    "presampling_fused": dict(do_light_presampling=1, light_presampling_number_of_subsets=8,
                              light_presampling_subset_size=32, do_fused_spatiotemporal=1,
                              spatial_number_of_passes=2),
}


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_restir_variants(mipt_mod, oracle_mod, torch_mod, variant):
    scene = mipt_mod.generate_scene(1, 64, 64)
    s = mipt_mod.default_render_settings(width=64, height=64, direct_light_sampling_strategy=5, nb_bounces=2)
    rs = mipt_mod.default_restir_di_settings(**VARIANTS[variant])
    render_both(mipt_mod, oracle_mod, scene, s, rs, 4)


@pytest.mark.parametrize("variant", ["power", "freeze"])
def test_fused_and_presampling_light_sampling(mipt_mod, oracle_mod, torch_mod, variant):
    """Presampled power-sampled lights into the fused pass; frozen random
    numbers (the fused pass then skips its temporal neighbour)."""
    scene = mipt_mod.generate_scene(1, 64, 64)
    s = mipt_mod.default_render_settings(width=64, height=64, direct_light_sampling_strategy=5, nb_bounces=2)
    if variant == "power":
        s.direct_light_sampling_base_strategy = 1
    else:
        s.freeze_random = 1
    rs = mipt_mod.default_restir_di_settings(do_fused_spatiotemporal=1, do_light_presampling=1,
                                             light_presampling_number_of_subsets=16,
                                             light_presampling_subset_size=64)
    render_both(mipt_mod, oracle_mod, scene, s, rs, 4, check_every_frame=True)


@pytest.mark.parametrize("variant", ["power", "freeze", "direct_only", "power_mis"])
def test_light_sampling_variants(mipt_mod, oracle_mod, torch_mod, variant):
    scene = mipt_mod.generate_scene(1, 64, 64)
    s = mipt_mod.default_render_settings(width=64, height=64, direct_light_sampling_strategy=5)
    if variant == "power":
        s.direct_light_sampling_base_strategy = 1
    elif variant == "freeze":
        s.freeze_random = 1
    elif variant == "direct_only":
        s.nb_bounces = 0
    else:
        s.direct_light_sampling_strategy = 3
        s.direct_light_sampling_base_strategy = 1
    render_both(mipt_mod, oracle_mod, scene, s, mipt_mod.default_restir_di_settings(), 4)


@pytest.mark.parametrize("guard", [0, 1])
def test_restir_config2_geometry(mipt_mod, oracle_mod, torch_mod, guard):
    """The 1M-triangle scene with 4096 power-sampled emitters and the
    ReSTIR DI parameters of the at-scale configuration (32 light + 1 BSDF
    candidates, temporal M-cap 20, 2 spatial passes of 5 neighbours in 16 px,
    5 bounces), at 240x135; with the reference's back-facing candidate weight
    (0/0) and with the guard."""
    scene = simplify_materials(mipt_mod.generate_scene(2, 240, 135))
    s = mipt_mod.default_render_settings(width=240, height=135, nb_bounces=5, direct_light_sampling_strategy=5,
                                         direct_light_sampling_base_strategy=1)
    rs = mipt_mod.default_restir_di_settings(number_of_initial_light_candidates=32, m_cap=20,
                                             spatial_number_of_passes=2, guard_backfacing_bsdf_candidates=guard)
    fb, R = render_both(mipt_mod, oracle_mod, scene, s, rs, 3)
    assert fb.mean() > 0.0


def test_restir_config2_fused_presampling(mipt_mod, oracle_mod, torch_mod):
    """The at-scale ReSTIR DI parameters with the fused spatiotemporal pass
    and the reference's default light presampling (128 subsets of 1024
    power-sampled lights, 8x8 pixel tiles per subset) on the 1M-triangle
    scene with 4096 emitters, at 240x135."""
    scene = simplify_materials(mipt_mod.generate_scene(2, 240, 135))
    s = mipt_mod.default_render_settings(width=240, height=135, nb_bounces=5, direct_light_sampling_strategy=5,
                                         direct_light_sampling_base_strategy=1)
    rs = mipt_mod.default_restir_di_settings(number_of_initial_light_candidates=32, m_cap=20,
                                             spatial_number_of_passes=2, do_fused_spatiotemporal=1,
                                             do_light_presampling=1)
    assert rs.light_presampling_number_of_subsets == 128 and rs.light_presampling_subset_size == 1024
    fb, R = render_both(mipt_mod, oracle_mod, scene, s, rs, 3)
    assert fb.mean() > 0.0


def test_restir_config2_layered_materials(mipt_mod, oracle_mod, torch_mod):
    """ReSTIR DI at the at-scale settings on the 1M-triangle scene with its
    layered principled materials (coat, sheen, dielectric specular,
    anisotropy), the glass lobe off."""
    scene = mipt_mod.generate_scene(2, 160, 90)
    for m in scene.materials:
        m.specular_transmission = 0.0
    s = mipt_mod.default_render_settings(width=160, height=90, nb_bounces=5, direct_light_sampling_strategy=5,
                                         direct_light_sampling_base_strategy=1)
    rs = mipt_mod.default_restir_di_settings(number_of_initial_light_candidates=32, m_cap=20,
                                             spatial_number_of_passes=2)
    fb, R = render_both(mipt_mod, oracle_mod, scene, s, rs, 2)
    assert fb.mean() > 0.0


def test_power_alias_table(mipt_mod, oracle_mod, torch_mod):
    scene = simplify_materials(mipt_mod.generate_scene(2, 64, 36))
    s = mipt_mod.default_render_settings(width=64, height=36, direct_light_sampling_base_strategy=1)
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    ctx.set_render_settings(s)
    E = scene.emissive.size
    probas, alias, total = ctx.power_alias_table(E)
    ctx.close()
    R = oracle_mod.Renderer(scene, s)
    cp, ca, ct = R.alias_table()
    assert E == 4096
    assert np.array_equal(probas.view(np.uint32), cp.view(np.uint32)) and np.array_equal(alias, ca) and total == ct
    # the table reproduces the power pmf
    pmf = probas.astype(np.float64).copy()
    np.add.at(pmf, alias, 1.0 - probas.astype(np.float64))
    pmf /= E
    assert abs(pmf.sum() - 1.0) < 1e-6 and pmf.min() > 0.0


def test_restir_pass_errors(mipt_mod, torch_mod):
    scene = mipt_mod.generate_scene(1, 32, 32)
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    ctx.set_render_settings(mipt_mod.default_render_settings(width=32, height=32))
    ctx.set_camera(scene.camera)
    with pytest.raises(mipt_mod.MiptError):  # strategy is MIS light + BSDF
        ctx.launch_pass(mipt_mod.MIPT_PASS_RESTIR_DI_INITIAL_CANDIDATES, 0, 1)
    ctx.set_render_settings(mipt_mod.default_render_settings(width=32, height=32, direct_light_sampling_strategy=5))
    with pytest.raises(mipt_mod.MiptError):  # spatial pass index beyond spatial_number_of_passes
        ctx.launch_pass(mipt_mod.MIPT_PASS_RESTIR_DI_SPATIAL_REUSE(1), 0, 1)
    with pytest.raises(mipt_mod.MiptError):
        ctx.set_restir_di_settings(mipt_mod.default_restir_di_settings(spatial_reuse_neighbor_count=33))
    for strategy in (0, 1, 2, 6):  # no light sampling, one light without MIS, BSDF only: not built; 6: none
        bad = mipt_mod.default_render_settings(width=32, height=32, direct_light_sampling_strategy=strategy)
        with pytest.raises(mipt_mod.MiptError):
            ctx.set_render_settings(bad)
    # the fused and presampling passes only run when their settings ask
    for p in (mipt_mod.MIPT_PASS_RESTIR_DI_SPATIOTEMPORAL_REUSE, mipt_mod.MIPT_PASS_RESTIR_DI_LIGHTS_PRESAMPLING):
        with pytest.raises(mipt_mod.MiptError):
            ctx.launch_pass(p, 0, 1)
    with pytest.raises(mipt_mod.MiptError):  # the fused pass reads last frame's reservoirs
        ctx.set_restir_di_settings(mipt_mod.default_restir_di_settings(do_fused_spatiotemporal=1,
                                                                       do_temporal_reuse_pass=0))
    with pytest.raises(mipt_mod.MiptError):
        ctx.set_restir_di_settings(mipt_mod.default_restir_di_settings(do_light_presampling=1,
                                                                       light_presampling_subset_size=0))
    ctx.set_restir_di_settings(mipt_mod.default_restir_di_settings(do_fused_spatiotemporal=1))
    with pytest.raises(mipt_mod.MiptError):  # spatial pass 0 is the fused pass
        ctx.launch_pass(mipt_mod.MIPT_PASS_RESTIR_DI_SPATIAL_REUSE(0), 0, 1)
    assert ctx.draws_per_sample() == 6
    # presampled records are read only after the presampling pass filled them
    ctx.set_restir_di_settings(mipt_mod.default_restir_di_settings(do_light_presampling=1))
    ctx.set_render_settings(mipt_mod.default_render_settings(width=32, height=32, direct_light_sampling_strategy=5))
    ctx.launch_pass(mipt_mod.MIPT_PASS_CAMERA_RAYS, 0, 1)
    with pytest.raises(mipt_mod.MiptError):
        ctx.launch_pass(mipt_mod.MIPT_PASS_RESTIR_DI_INITIAL_CANDIDATES, 0, 1)
    ctx.launch_pass(mipt_mod.MIPT_PASS_RESTIR_DI_LIGHTS_PRESAMPLING, 0, 1)
    ctx.launch_pass(mipt_mod.MIPT_PASS_RESTIR_DI_INITIAL_CANDIDATES, 0, 1)
    ctx.close()


def test_restir_temporal_surface_that_missed(mipt_mod, oracle_mod, torch_mod, open_back_box_scene):
    """Regression: pixels of the open-back box go hit -> miss -> hit under
    camera jitter, so the temporal pass evaluates a previous-frame surface
    without a primary hit (prim -1, material of an older hit). Reading the
    material index of primitive -1 there made the illegal memory access of
    ReSTIR DI at 1920x1080; the surface now keeps the empty volume stack
    (RayVolumeState.h:37-39). Every frame bit-exact against the oracle."""
    scene = open_back_box_scene(mipt_mod, 32, 32)
    s = mipt_mod.default_render_settings(width=32, height=32, nb_bounces=1, direct_light_sampling_strategy=5)
    render_both(mipt_mod, oracle_mod, scene, s, mipt_mod.default_restir_di_settings(), 5, check_every_frame=True)

