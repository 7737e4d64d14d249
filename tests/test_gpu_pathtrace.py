"""GPU parity of the path-tracing passes (CameraRays + MegaKernel semantics on
the wavefront kernels) against the oracle's per-pixel megakernel restatement,
through the C ABI.

Bar (the tolerance the plain method is held to): first-hit primitive IDs
bit-exact; per-pixel radiance within 1e-4 relative on >= 99.9 % of pixels
(per channel |gpu - cpu| <= 1e-4 * max(|gpu|, |cpu|) + 1e-7). Pixels outside
it would be FP32-rounding branch divergences and are counted and reported.
Both sides evaluate sin/cos/pow by the same fixed operation sequence, so the
tests also assert the stronger property that every pixel is bit-identical.
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


def _render_gpu(mipt_mod, scene, settings, spp):
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    ctx.set_render_settings(settings)
    ctx.set_camera(scene.camera)
    for sn in range(spp):
        ctx.render_sample(sn)
    fb, fb64 = ctx.read_framebuffer(fp64=True)
    prims = ctx.read_gbuffer_prims()
    rejected = ctx.rejected_samples()
    ctx.close()
    return fb, fb64, prims, rejected


def _render_cpu(mipt_mod, oracle_mod, scene, settings, spp):
    R = oracle_mod.Renderer(scene, settings)
    for sn in range(spp):
        R.render_sample(sn, mipt_mod.host_random_number(2 * sn), mipt_mod.host_random_number(2 * sn + 1))
    return R.fb, R.fb64, R.gb_prims, R.rejected


def _compare(gpu, cpu):
    g = gpu.reshape(-1, 3).astype(np.float64)
    c = cpu.reshape(-1, 3).astype(np.float64)
    ok = np.abs(g - c) <= REL_TOL * np.maximum(np.abs(g), np.abs(c)) + ABS_TOL
    px_ok = ok.all(axis=1)
    return px_ok.mean(), int((~px_ok).sum())


def _check(mipt_mod, oracle_mod, scene, settings, spp):
    fb, fb64, prims, rej = _render_gpu(mipt_mod, scene, settings, spp)
    cfb, cfb64, cprims, crej = _render_cpu(mipt_mod, oracle_mod, scene, settings, spp)
    assert np.array_equal(prims, cprims), "first-hit primitive IDs differ"
    frac, diverged = _compare(fb, cfb)
    frac64, diverged64 = _compare(fb64, cfb64)
    identical = int((fb.reshape(-1, 3) == cfb.reshape(-1, 3)).all(axis=1).sum())
    print(f"\n{settings.width}x{settings.height} {spp} spp: {frac * 100:.4f} % pixels within 1e-4 "
          f"({diverged} divergent, {identical} of {fb.size // 3} bit-identical), fp64 {frac64 * 100:.4f} % "
          f"({diverged64}); rejected gpu {rej} cpu {crej}")
    assert frac >= MIN_PASS
    assert frac64 >= MIN_PASS
    assert rej == crej
    # Stronger than the bar: sin/cos/pow are the same fixed sequences on both
    # sides, so every pixel agrees bit for bit.
    assert identical == fb.size // 3
    assert np.array_equal(fb64, cfb64)
    # the image is not trivially black
    assert fb.mean() > 0.01 * spp
    return fb, cfb


def test_config1_small(mipt_mod, oracle_mod, torch_mod):
    scene = mipt_mod.generate_scene(1, 128, 128)
    s = mipt_mod.default_render_settings(width=128, height=128)
    _check(mipt_mod, oracle_mod, scene, s, 4)


@pytest.mark.parametrize("variant", ["no_rr", "no_direct_at_bounce0", "freeze", "no_regularization", "one_bounce"])
def test_config1_settings_variants(mipt_mod, oracle_mod, torch_mod, variant):
    scene = mipt_mod.generate_scene(1, 96, 96)
    s = mipt_mod.default_render_settings(width=96, height=96)
    if variant == "no_rr":
        s.do_russian_roulette = 0
    elif variant == "no_direct_at_bounce0":
        s.enable_direct = 0
    elif variant == "freeze":
        s.freeze_random = 1
    elif variant == "no_regularization":
        s.do_microfacet_regularization = 0
    else:
        s.nb_bounces = 1
    _check(mipt_mod, oracle_mod, scene, s, 3)


def test_config1_full(mipt_mod, oracle_mod, torch_mod):
    """The plain-method configuration: 512x512, 16 one-sample passes,
    4 bounces, uniform light sampling NEE + MIS with BSDF sampling."""
    scene = mipt_mod.generate_scene(1, 512, 512)
    s = mipt_mod.default_render_settings(width=512, height=512, nb_bounces=4)
    fb, cfb = _check(mipt_mod, oracle_mod, scene, s, 16)
    mean = (fb.reshape(-1, 3) / 16).mean(axis=0)
    assert (mean > 0.05).all() and (mean < 5).all()


def _layered(m, rng, smooth=False):
    """Random coat / sheen / dielectric specular / metal / anisotropy
    parameters (no glass, diffuse transmission or thin film)."""
    m.metallic = float(rng.uniform() < 0.3) * rng.uniform()
    m.roughness = 0.0 if smooth else rng.uniform(0.05, 1.0)
    m.specular = rng.uniform()
    m.specular_tint = rng.uniform()
    for k in range(3):
        m.specular_color[k] = rng.uniform(0.3, 1.0)
        m.sheen_color[k] = rng.uniform(0.3, 1.0)
    m.coat = rng.uniform()
    m.coat_roughness = 0.0 if smooth else rng.uniform()
    m.coat_ior = rng.uniform(1.2, 1.8)
    m.sheen = rng.uniform()
    m.sheen_roughness = rng.uniform(0.05, 1.0)
    m.anisotropy = rng.uniform()
    m.anisotropy_rotation = rng.uniform()
    m.ior = rng.uniform(1.1, 2.0)
    m.second_roughness_weight = rng.uniform() * 0.5
    m.specular_transmission = 0.0
    m.diffuse_transmission = 0.0
    m.thin_film = 0.0


def test_config1_layered_materials(mipt_mod, oracle_mod, torch_mod):
    """Config-1 geometry with every non-emissive material given random coat,
    sheen, dielectric specular, metal and anisotropy parameters; two of them
    perfectly smooth (delta coat / specular / metal lobes) and one with an
    absorbing coat medium."""
    scene = mipt_mod.generate_scene(1, 96, 96)
    rng = np.random.default_rng(21)
    k = 0
    for m in scene.materials:
        if max(m.emission) > 0:
            continue
        _layered(m, rng, smooth=k in (3, 7))
        if k == 5:
            m.coat = 1.0
            for c, a in enumerate((0.4, 0.7, 0.9)):
                m.coat_medium_absorption[c] = a
        k += 1
    s = mipt_mod.default_render_settings(width=96, height=96)
    _check(mipt_mod, oracle_mod, scene, s, 3)


def test_config2_layered_materials(mipt_mod, oracle_mod, torch_mod):
    """The 1M-triangle scene with its principled materials (coat, sheen,
    dielectric specular, anisotropy drawn per object), the glass lobe off;
    uniform NEE + MIS, 5 bounces."""
    scene = mipt_mod.generate_scene(2, 160, 90)
    for m in scene.materials:
        m.specular_transmission = 0.0
    s = mipt_mod.default_render_settings(width=160, height=90, nb_bounces=5)
    _check(mipt_mod, oracle_mod, scene, s, 2)


def test_scene_without_emitters_is_black(mipt_mod, oracle_mod, torch_mod):
    scene = mipt_mod.generate_scene(1, 64, 64)
    scene.emissive = np.zeros(0, np.int32)
    # drop the emission of the two light materials too
    for i in range(len(scene.materials)):
        for k in range(3):
            scene.materials[i].emission[k] = 0.0
    s = mipt_mod.default_render_settings(width=64, height=64)
    fb, fb64, prims, rej = _render_gpu(mipt_mod, scene, s, 2)
    assert fb.max() == 0.0 and rej == 0
    cfb, _, cprims, _ = _render_cpu(mipt_mod, oracle_mod, scene, s, 2)
    assert np.array_equal(prims, cprims) and cfb.max() == 0.0


def test_framebuffer_is_a_sum(mipt_mod, torch_mod):
    """PathTracing.h:176-217: sample 0 overwrites, later samples add."""
    scene = mipt_mod.generate_scene(1, 64, 64)
    s = mipt_mod.default_render_settings(width=64, height=64)
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    ctx.set_render_settings(s)
    ctx.set_camera(scene.camera)
    ctx.render_sample(0)
    a = ctx.read_framebuffer()
    ctx.render_sample(1)
    b = ctx.read_framebuffer()
    ctx.render_sample(0)  # restarting at sample 0 overwrites the sum
    c = ctx.read_framebuffer()
    assert np.array_equal(a, c)
    assert (b >= a - 1e-6).all() and b.sum() > a.sum()
    ctx.close()


@pytest.mark.parametrize("scene_kind", ["config1", "nested_glass", "cache"])
def test_shading_order_does_not_change_the_image(mipt_mod, torch_mod, scene_kind):
    """The path pass shades the live paths in list order or sorted by
    material lobe class (and stored volume state); each path's computation
    depends only on its own state, so the framebuffer (FP32 and FP64 sums)
    is bit-identical in every order: the config-1 scene (Lambertian and
    metal classes) and the nested-glass scene (glass, diffuse transmission,
    coat, paths inside media). With the radiance cache (the sorted list is
    then the training paths' list past the query bounce) the paths are the
    same, but the network's gradients are FP32 atomic sums whose order
    follows the records' order, so the trained weights and the images agree
    to rounding only: mean relative difference below 1e-3."""
    from test_gpu_glass import nested_glass_scene

    w = h = 128
    scene = nested_glass_scene(mipt_mod, w, h) if scene_kind == "nested_glass" else \
        mipt_mod.generate_scene(1, w, h)
    cache = scene_kind == "cache"
    out = {}
    for order in (mipt_mod.MIPT_SHADE_ORDER_LIVE, mipt_mod.MIPT_SHADE_ORDER_MATERIAL,
                  mipt_mod.MIPT_SHADE_ORDER_AUTO):
        ctx = mipt_mod.Context(0)
        ctx.set_scene(scene)
        ctx.build_bvh()
        ctx.set_shading_order(order)
        if cache:
            ctx.set_radiance_cache(mipt_mod.default_radiance_cache_settings())
        ctx.set_render_settings(mipt_mod.default_render_settings(width=w, height=h, nb_bounces=6,
                                                                 accumulate_fp64=1, radiance_cache=int(cache)))
        ctx.set_camera(scene.camera)
        for s in range(3):
            ctx.render_sample(s)
        fb, fb64 = ctx.read_framebuffer(fp64=True)
        out[order] = (fb.tobytes(), fb64.tobytes()) if not cache else fb64.astype(np.float64)
        ctx.close()
    if cache:
        for o in (1, 2):
            rel = np.abs(out[o] - out[0]) / (np.abs(out[0]) + 1e-3)
            assert rel.mean() < 1e-3, (o, rel.mean())
    else:
        assert out[0] == out[1] == out[2]
    with pytest.raises(mipt_mod.MiptError):
        ctx = mipt_mod.Context(0)
        try:
            ctx.set_shading_order(3)
        finally:
            ctx.close()


def smooth_sphere_scene(mipt_mod, w, h, config=1):
    """The config-1 (or config-2) scene with vertex normals on every
    icosphere: normalize(vertex - sphere centre); the walls, the terrain and
    the lights keep geometric normals (no flag on their vertices)."""
    scene = mipt_mod.generate_scene(config, w, h)
    mi = scene.material_indices
    vn = np.zeros_like(scene.vertices)
    has = np.zeros(scene.vertices.shape[0], np.uint8)
    per_sphere = 5120 if config == 1 else 1280
    for k in range(len(scene.materials)):
        sel = mi == k
        if sel.sum() != per_sphere:
            continue
        used = np.unique(scene.indices[sel])
        d = scene.vertices[used].astype(np.float64) - scene.vertices[used].astype(np.float64).mean(axis=0)
        vn[used] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
        has[used] = 1
    assert has.sum() > 0
    return mipt_mod.Scene(scene.vertices, scene.indices, scene.material_indices, scene.materials, scene.emissive,
                          scene.camera, w, h, vertex_normals=vn, has_vertex_normals=has)


def test_config1_smooth_normals(mipt_mod, oracle_mod, torch_mod):
    """Smooth shading normals (get_shading_normal, Intersect.h:114-133) on
    the spheres: bit-exact against the oracle; the image differs from the
    flat-shaded one, and do_normal_mapping = 0 gives the flat-shaded image."""
    w = h = 128
    scene = smooth_sphere_scene(mipt_mod, w, h)
    s = mipt_mod.default_render_settings(width=w, height=h)
    fb, _ = _check(mipt_mod, oracle_mod, scene, s, 4)
    flat = mipt_mod.Scene(scene.vertices, scene.indices, scene.material_indices, scene.materials, scene.emissive,
                          scene.camera, w, h)
    fb_flat = _render_gpu(mipt_mod, flat, s, 4)[0]
    off = _render_gpu(mipt_mod, scene, mipt_mod.default_render_settings(width=w, height=h, do_normal_mapping=0), 4)[0]
    assert off.tobytes() == fb_flat.tobytes()
    changed = (fb.reshape(-1, 3) != fb_flat.reshape(-1, 3)).any(axis=1).mean()
    assert changed > 0.05


def test_restir_smooth_normals(mipt_mod, oracle_mod, torch_mod):
    """ReSTIR DI surfaces read the G-buffer's shading normal: the
    config-2 geometry with smooth spheres at 160x90, at-scale ReSTIR
    settings, bit-exact against the oracle over 3 frames."""
    from test_gpu_restir import render_both, simplify_materials

    scene = simplify_materials(smooth_sphere_scene(mipt_mod, 160, 90, config=2))
    s = mipt_mod.default_render_settings(width=160, height=90, nb_bounces=3, direct_light_sampling_strategy=5,
                                         direct_light_sampling_base_strategy=1)
    rs = mipt_mod.default_restir_di_settings(number_of_initial_light_candidates=32, m_cap=20,
                                             spatial_number_of_passes=2)
    fb, R = render_both(mipt_mod, oracle_mod, scene, s, rs, 3, check_every_frame=True)
    assert fb.mean() > 0.0


def test_update_materials_without_rebuild(mipt_mod, torch_mod):
    """GPURenderer::update_all_materials / update_one_material: new materials
    on a context that has rendered, without a BVH rebuild, render exactly
    what a fresh context built with them renders; a glass material switches
    the shading order's volume handling on; a wrong count or index is
    refused and leaves the materials as they were."""
    import ctypes as C

    w = h = 96
    scene = mipt_mod.generate_scene(1, w, h)
    s = mipt_mod.default_render_settings(width=w, height=h)
    changed = (type(scene.materials[0]) * len(scene.materials))(*scene.materials)
    sphere = [k for k in range(len(changed)) if (scene.material_indices == k).sum() == 5120]
    changed[sphere[0]].metallic = 1.0
    changed[sphere[0]].roughness = 0.2
    changed[sphere[1]].specular_transmission = 1.0
    changed[sphere[1]].metallic = 0.0
    changed[sphere[1]].ior = 1.5
    changed[sphere[2]].base_color[0] = 0.05

    def fresh(mats):
        ctx = mipt_mod.Context(0)
        ctx.set_scene(mipt_mod.Scene(scene.vertices, scene.indices, scene.material_indices, mats, scene.emissive,
                                     scene.camera, w, h))
        ctx.build_bvh()
        ctx.set_render_settings(s)
        ctx.set_camera(scene.camera)
        return ctx

    def render(ctx):
        for sn in range(3):
            ctx.render_sample(sn)
        return ctx.read_framebuffer().tobytes()

    ref = fresh(changed)
    want = render(ref)
    ref.close()
    ctx = fresh(scene.materials)
    before = render(ctx)
    ctx.update_materials(changed)
    assert render(ctx) == want
    ctx.update_materials(scene.materials)
    assert render(ctx) == before
    for k in (sphere[0], sphere[1], sphere[2]):
        ctx.update_material(k, changed[k])
    assert render(ctx) == want
    with pytest.raises(mipt_mod.MiptError):
        ctx.update_material(len(changed), changed[0])
    short = (type(scene.materials[0]) * 3)(*list(scene.materials)[:3])
    with pytest.raises(mipt_mod.MiptError):
        ctx.update_materials(short)
    assert render(ctx) == want
    ctx.close()
