"""GPU parity of the ambient light (WorldSettings, Envmap.h) through the C ABI,
against the oracle's restatement:

- the packed R9G9B9E5 texels, the alias table and the luminance sum of
  mipt_set_envmap are the oracle's bit for bit (the GPU packs with an exact
  power-of-two scale and round-half-even, the oracle with the reference's
  bias addition);
- frames with the envmap sampled at every vertex (light sample + BSDF sample
  with MIS, two extra shadow rays per vertex) and seen by camera rays, with
  emissive triangles or without, rotated, scaled, under a uniform ambient
  colour or none: first-hit IDs and every pixel bit-identical (the bar is
  1e-4 relative on >= 99.9 % of pixels);
- the same in a batch of views and under graph replay;
- ReSTIR DI with envmap candidates, reservoirs and frames bit-identical;
- the combination this build does not implement (the radiance cache with
  an envmap) is refused.
"""
import numpy as np
import pytest

from conftest import sky_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


@pytest.fixture(scope="module")
def envmap(mipt_mod):
    return mipt_mod.generate_envmap(256, 128)


def rotation_y(deg):
    a = np.radians(deg)
    rows = np.array([[np.cos(a), 0, -np.sin(a)], [0, 1, 0], [np.sin(a), 0, np.cos(a)]], np.float32)
    return rows.T.copy(), rows.copy()


def _ctx(mipt_mod, scene, settings, world, env):
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    if env is not None:
        ctx.set_envmap(env)
    if world is not None:
        ctx.set_world_settings(world)
    ctx.set_render_settings(settings)
    ctx.set_camera(scene.camera)
    return ctx


def render_both(mipt_mod, oracle_mod, scene, settings, world, env, spp):
    ctx = _ctx(mipt_mod, scene, settings, world, env)
    for sn in range(spp):
        ctx.render_sample(sn)
    fb, fb64 = ctx.read_framebuffer(fp64=True)
    prims = ctx.read_gbuffer_prims()
    rej = ctx.rejected_samples()
    ctx.close()
    R = oracle_mod.Renderer(scene, settings, world=world, envmap=env)
    for sn in range(spp):
        R.render_sample(sn, mipt_mod.host_random_number(2 * sn), mipt_mod.host_random_number(2 * sn + 1))
    assert np.array_equal(prims, R.gb_prims), "first-hit IDs differ"
    identical = int((fb.reshape(-1, 3) == R.fb.reshape(-1, 3)).all(axis=1).sum())
    print(f"\n{settings.width}x{settings.height} {spp} spp: {identical} of {fb.size // 3} pixels bit-identical, "
          f"{int((prims < 0).sum())} background; rejected gpu {rej} cpu {R.rejected}")
    assert identical == fb.size // 3
    assert np.array_equal(fb64, R.fb64)
    assert rej == R.rejected
    return fb, R


@pytest.mark.parametrize("case", ["generated", "lognormal_rgba"])
def test_envmap_tables_equal_oracle(mipt_mod, oracle_mod, torch_mod, case):
    if case == "generated":
        env = mipt_mod.generate_envmap(512, 256)
    else:
        rng = np.random.default_rng(11)
        env = np.exp(rng.normal(0, 2, (96, 160, 4))).astype(np.float32)
    scene = sky_scene(mipt_mod, 16, 16)
    ctx = _ctx(mipt_mod, scene, mipt_mod.default_render_settings(width=16, height=16), None, env)
    assert ctx.world_settings().ambient_light_type == mipt_mod.MIPT_AMBIENT_ENVMAP
    n = env.shape[0] * env.shape[1]
    probas, alias, total, texels = ctx.envmap_tables(n)
    ctx.close()
    R = oracle_mod.Renderer(scene, mipt_mod.default_render_settings(width=16, height=16), envmap=env)
    cp, ca, ct, ctex = R.env_tables()
    assert np.array_equal(texels, ctex), int((texels != ctex).sum())
    assert np.array_equal(probas.view(np.uint32), cp.view(np.uint32)) and np.array_equal(alias, ca)
    assert total == ct


VARIANTS = {
    "default": dict(),
    "scaled_background": dict(world=dict(envmap_intensity=2.5, envmap_scale_background_intensity=1)),
    "rotated": dict(world="rotated"),
    "deep_paths": dict(settings=dict(nb_bounces=7)),
    "no_emitters": dict(lights=False),
    "power": dict(settings=dict(direct_light_sampling_base_strategy=1)),
    "freeze": dict(settings=dict(freeze_random=1)),
    "uniform": dict(world=dict(ambient_light_type=1, uniform_light_color=[0.3, 0.5, 0.8]), env=False),
    "none": dict(world=dict(ambient_light_type=0), env=False),
    "intensity_zero": dict(world=dict(envmap_intensity=0.0)),
}


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_envmap_render(mipt_mod, oracle_mod, torch_mod, envmap, variant):
    v = VARIANTS[variant]
    scene = sky_scene(mipt_mod, 64, 64, lights=v.get("lights", True))
    settings = mipt_mod.default_render_settings(width=64, height=64, nb_bounces=3)
    for k, val in v.get("settings", {}).items():
        setattr(settings, k, val)
    env = envmap if v.get("env", True) else None
    w = v.get("world", {})
    if w == "rotated":
        to_world, to_env = rotation_y(60.0)
        world = mipt_mod.default_world_settings(ambient_light_type=2, envmap_to_world=to_world,
                                                world_to_envmap=to_env)
    else:
        w = dict(w)
        w.setdefault("ambient_light_type", 2 if env is not None else 1)
        world = mipt_mod.default_world_settings(**w)
    fb, R = render_both(mipt_mod, oracle_mod, scene, settings, world, env, 4)
    if variant != "none" and variant != "intensity_zero":
        assert fb.mean() > 0.01


def test_envmap_config2_layered(mipt_mod, oracle_mod, torch_mod, envmap):
    """The 1M-triangle scene with all its layered materials (glass,
    dispersion, coat, sheen, thin film) under the envmap, 5 bounces."""
    scene = mipt_mod.generate_scene(2, 160, 90)
    settings = mipt_mod.default_render_settings(width=160, height=90, nb_bounces=5,
                                                direct_light_sampling_base_strategy=1)
    world = mipt_mod.default_world_settings(ambient_light_type=2, envmap_intensity=0.5)
    render_both(mipt_mod, oracle_mod, scene, settings, world, envmap, 2)


def test_envmap_views_and_graph(mipt_mod, torch_mod, envmap):
    """A batch of views renders what each view renders alone, and graph
    replays equal eager samples, with the envmap on."""
    scene = sky_scene(mipt_mod, 64, 64)
    views = mipt_mod.generate_views(4)
    for i, v in enumerate(views):
        v.camera = scene.camera
        v.width, v.height = 40 + 8 * i, 32 + 4 * i
    world = mipt_mod.default_world_settings(ambient_light_type=2)
    s = mipt_mod.default_render_settings(width=64, height=64, nb_bounces=3)
    batch = _ctx(mipt_mod, scene, s, world, envmap)
    batch.set_views(views)
    eager = _ctx(mipt_mod, scene, s, world, envmap)
    eager.set_views(views)
    batch.set_graph_capture(True)
    for sn in range(4):
        batch.render_sample(sn)
        eager.render_sample(sn)
        torch_mod.cuda.synchronize()
        assert batch.read_framebuffer().tobytes() == eager.read_framebuffer().tobytes(), sn
    assert batch.graph_replays() == 3
    fb = batch.read_framebuffer()
    batch.close()
    eager.close()
    off = 0
    for v in views:
        one = _ctx(mipt_mod, scene, mipt_mod.default_render_settings(width=v.width, height=v.height, nb_bounces=3),
                   world, envmap)
        one.set_camera(v.camera)
        for sn in range(4):
            one.render_sample(sn)
        f1 = one.read_framebuffer()
        one.close()
        assert fb[off:off + f1.size].tobytes() == f1.tobytes()
        off += f1.size


def render_restir_both(mipt_mod, oracle_mod, scene, settings, restir, world, env, frames):
    ctx = _ctx(mipt_mod, scene, settings, world, env)
    ctx.set_restir_di_settings(restir)
    ctx.set_render_settings(settings)
    ctx.set_camera(scene.camera)
    draws = ctx.draws_per_sample()
    R = oracle_mod.Renderer(scene, settings, restir, world=world, envmap=env)
    for f in range(frames):
        ctx.render_sample(f)
        R.render_frame(f, [mipt_mod.host_random_number(f * draws + k) for k in range(draws)])
    assert np.array_equal(ctx.read_gbuffer_prims(), R.gb_prims), "first-hit IDs differ"
    for which in (0, 1, 2, 3):
        g, c = ctx.read_reservoirs(which), R.read_reservoirs(which)
        same = (g.view(np.uint32).reshape(len(g), -1) == c.view(np.uint32).reshape(len(c), -1)).all(axis=1)
        assert same.all(), (which, np.nonzero(~same)[0][:5])
    fb, fb64 = ctx.read_framebuffer(fp64=True)
    rej = ctx.rejected_samples()
    ctx.close()
    identical = int((fb.reshape(-1, 3) == R.fb.reshape(-1, 3)).all(axis=1).sum())
    envs = int(((R.read_reservoirs(3)["flags"] & 1) == 1).sum())
    print(f"\n{settings.width}x{settings.height} x{frames}: {identical} of {fb.size // 3} pixels bit-identical; "
          f"{envs} envmap reservoirs; rejected gpu {rej} cpu {R.rejected}")
    assert identical == fb.size // 3 and np.array_equal(fb64, R.fb64) and rej == R.rejected
    return fb, R, envs


RESTIR_VARIANTS = {
    "default": dict(),
    "initial_only": dict(do_temporal_reuse_pass=0, do_spatial_reuse_pass=0),
    "two_spatial_prob_02": dict(spatial_number_of_passes=2, envmap_candidate_probability=0.2),
    "fused_presampled": dict(do_fused_spatiotemporal=1, do_light_presampling=1,
                             light_presampling_number_of_subsets=16, light_presampling_subset_size=64),
    "envmap_only": dict(lights=False),
    "rotated": dict(rotated=True),
}


@pytest.mark.parametrize("variant", list(RESTIR_VARIANTS))
def test_restir_with_envmap(mipt_mod, oracle_mod, torch_mod, envmap, variant):
    """ReSTIR DI with envmap candidates (light candidates and presampled
    lights drawn from the envmap with the envmap candidate probability, BSDF
    candidates that leave the scene) through every pass: reservoirs and
    frames bit-identical."""
    kw = dict(RESTIR_VARIANTS[variant])
    lights = kw.pop("lights", True)
    rotated = kw.pop("rotated", False)
    scene = sky_scene(mipt_mod, 64, 64, lights=lights)
    settings = mipt_mod.default_render_settings(width=64, height=64, nb_bounces=2, direct_light_sampling_strategy=5)
    restir = mipt_mod.default_restir_di_settings(**kw)
    if rotated:
        to_world, to_env = rotation_y(60.0)
        world = mipt_mod.default_world_settings(ambient_light_type=2, envmap_to_world=to_world,
                                                world_to_envmap=to_env)
    else:
        world = mipt_mod.default_world_settings(ambient_light_type=2)
    fb, R, envs = render_restir_both(mipt_mod, oracle_mod, scene, settings, restir, world, envmap, 4)
    assert envs > 0 and fb.mean() > 0.01


def test_restir_with_envmap_config2(mipt_mod, oracle_mod, torch_mod, envmap):
    """The at-scale ReSTIR DI parameters (32 + 1 candidates, M-cap 20, two
    spatial passes, fused pass, default presampling) on the 1M-triangle scene
    with its 4096 emitters and the envmap, 240x135."""
    scene = mipt_mod.generate_scene(2, 240, 135)
    for m in scene.materials:
        m.specular_transmission = 0.0
    settings = mipt_mod.default_render_settings(width=240, height=135, nb_bounces=5, direct_light_sampling_strategy=5,
                                                direct_light_sampling_base_strategy=1)
    restir = mipt_mod.default_restir_di_settings(number_of_initial_light_candidates=32, m_cap=20,
                                                 spatial_number_of_passes=2, do_fused_spatiotemporal=1,
                                                 do_light_presampling=1)
    world = mipt_mod.default_world_settings(ambient_light_type=2, envmap_intensity=0.5)
    render_restir_both(mipt_mod, oracle_mod, scene, settings, restir, world, envmap, 3)


def test_radiance_cache_with_envmap(mipt_mod, torch_mod, envmap):
    """The learned radiance cache under the envmap (sky scene, 128 x 128, 4
    bounces): queries and training records are made, no target is dropped,
    and after warm-up its 16-spp image has the mean of an FP64 1024-spp
    plain render (2 %) and an error within 1.25x the plain form's."""
    w = h = 128
    scene = sky_scene(mipt_mod, w, h)
    world = mipt_mod.default_world_settings(ambient_light_type=2, envmap_intensity=0.5)

    def ctx_for(cache):
        s = mipt_mod.default_render_settings(width=w, height=h, nb_bounces=4, accumulate_fp64=1,
                                             radiance_cache=1 if cache else 0)
        ctx = mipt_mod.Context(0)
        ctx.set_scene(scene)
        ctx.build_bvh()
        ctx.set_envmap(envmap)
        ctx.set_world_settings(world)
        if cache:
            ctx.set_radiance_cache(mipt_mod.default_radiance_cache_settings(batch_size=1 << 14))
        ctx.set_render_settings(s)
        ctx.set_camera(scene.camera)
        return ctx

    def render(ctx, spp, first=1):
        for i in range(spp):
            ctx.render_sample(first + i - 1 if i else 0)  # sample 0 starts the sum
        torch_mod.cuda.synchronize()
        return ctx.read_framebuffer(fp64=True)[1] / spp

    ref_ctx = ctx_for(False)
    ref = render(ref_ctx, 1024, first=5000)
    plain = render(ref_ctx, 16)  # a fresh 16-spp accumulation on the same context
    ref_ctx.close()
    cache = ctx_for(True)
    cache.render_sample(0)
    torch_mod.cuda.synchronize()
    queries, train, dropped, loss = cache.cache_stats()
    assert dropped == 0 and queries > 0 and train > 0 and np.isfinite(loss)
    for i in range(64):
        cache.cache_warmup(10000 + i)
    img = render(cache, 16)
    cache.close()

    def relmse(x, r):
        return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))

    e_plain, e_cache = relmse(plain, ref), relmse(img, ref)
    print(f"\nenvmap + cache: relMSE plain {e_plain:.4g} cache {e_cache:.4g}; mean plain {plain.mean():.4f} "
          f"cache {img.mean():.4f} ref {ref.mean():.4f}")
    assert np.isfinite(img).all() and (img >= 0).all()
    # under an open sky most light arrives at the first two vertices, whose
    # variance the cache does not touch: no gain to claim here (measured
    # 0.00938 against 0.00913), but no loss either, and the same mean
    assert e_cache < 1.25 * e_plain
    assert abs(img.mean() - ref.mean()) < 0.02 * ref.mean()


def test_envmap_errors(mipt_mod, torch_mod, envmap):
    scene = sky_scene(mipt_mod, 32, 32)
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    with pytest.raises(mipt_mod.MiptError):  # no envmap yet
        ctx.set_world_settings(mipt_mod.default_world_settings(ambient_light_type=2))
    with pytest.raises(mipt_mod.MiptError):  # two channels
        ctx.set_envmap(np.ones((8, 16, 2), np.float32))
    ctx.set_envmap(envmap)
    with pytest.raises(mipt_mod.MiptError):  # a probability outside [0, 1]
        ctx.set_restir_di_settings(mipt_mod.default_restir_di_settings(envmap_candidate_probability=1.5))
    ctx.set_render_settings(mipt_mod.default_render_settings(width=32, height=32))
    ctx.set_camera(scene.camera)
    ctx.render_sample(0)
    # an empty envmap turns the ambient light back to uniform
    ctx.set_envmap(None)
    assert ctx.world_settings().ambient_light_type == mipt_mod.MIPT_AMBIENT_UNIFORM
    ctx.render_sample(1)
    ctx.close()
