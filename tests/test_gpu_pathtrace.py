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
    metal classes), the nested-glass scene (glass, diffuse transmission,
    coat, paths inside media) and the config-1 scene with the radiance
    cache (the sorted list is then the training paths' list past the query
    bounce)."""
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
        out[order] = (fb.tobytes(), fb64.tobytes())
        ctx.close()
    assert out[0] == out[1] == out[2]
    with pytest.raises(mipt_mod.MiptError):
        ctx = mipt_mod.Context(0)
        try:
            ctx.set_shading_order(3)
        finally:
            ctx.close()
