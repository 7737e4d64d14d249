"""The full stack at scale (configuration 5): the 10,479,104-triangle scene
with 16,384 emitters and every Principled layer, under a procedural sun + sky
envmap, through the C ABI:

- the SAH tree of the 10M triangles gives, at 480x270, first-hit IDs equal
  to the oracle's camera pass on every pixel;
- ReSTIR DI (32 + 1 candidates, temporal M-cap 20, 2 spatial passes, envmap
  candidates) with 8 bounces renders 2 frames at 480x270: on a 24x24 tile
  every reservoir buffer and the framebuffer bit for bit against the oracle's
  windowed restatement (the oracle computes only the pixels the tile depends
  on). ReSTIR DI does not take dispersive glass yet, so the dispersion of
  transmissive materials is set to 0 here (the other layers stay).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, H = 480, 270
TILE = 24


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


@pytest.fixture(scope="module")
def scene5(mipt_mod):
    scene = mipt_mod.generate_scene(5, W, H)
    for m in scene.materials:
        if m.dispersion_scale > 0.0 and m.specular_transmission > 0.0:
            m.dispersion_scale = 0.0
    return scene


def test_config5_first_hits(mipt_mod, oracle_mod, torch_mod, scene5):
    assert scene5.triangle_count == 10479104 and scene5.emissive.size == 16384
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene5)
    ctx.build_bvh(mipt_mod.MIPT_BVH_SAH_BINNED)
    info = ctx.bvh_info()
    s = mipt_mod.default_render_settings(width=W, height=H)
    ctx.set_render_settings(s)
    ctx.set_camera(scene5.camera)
    ctx.launch_pass(mipt_mod.MIPT_PASS_CAMERA_RAYS, 0, mipt_mod.host_random_number(0))
    torch_mod.cuda.synchronize()
    prims = ctx.read_gbuffer_prims()
    ctx.close()
    print(f"\nSAH build of {scene5.triangle_count} triangles: {info.build_ms:.1f} ms, {info.node_count} nodes, "
          f"SAH cost {info.sah_cost:.1f}; {int((prims >= 0).sum())} of {W * H} pixels hit")
    R = oracle_mod.Renderer(scene5, s)
    R.camera_pass(0, mipt_mod.host_random_number(0))
    assert np.array_equal(prims, R.gb_prims)
    assert (prims >= 0).mean() > 0.5


def test_config5_restir_envmap_tile(mipt_mod, oracle_mod, torch_mod, scene5):
    env = mipt_mod.generate_envmap(512, 256)
    world = mipt_mod.default_world_settings(ambient_light_type=mipt_mod.MIPT_AMBIENT_ENVMAP)
    s = mipt_mod.default_render_settings(width=W, height=H, nb_bounces=8, direct_light_sampling_strategy=5,
                                         direct_light_sampling_base_strategy=1)
    rs = mipt_mod.default_restir_di_settings(number_of_initial_light_candidates=32, m_cap=20,
                                             spatial_number_of_passes=2)
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene5)
    ctx.build_bvh(mipt_mod.MIPT_BVH_SAH_BINNED)
    ctx.set_envmap(env)
    ctx.set_world_settings(world)
    ctx.set_restir_di_settings(rs)
    ctx.set_render_settings(s)
    ctx.set_camera(scene5.camera)
    draws = ctx.draws_per_sample()
    frames = 2
    for f in range(frames):
        print(f"config 5 ReSTIR + envmap: frame {f}", flush=True)
        ctx.render_sample(f)
        torch_mod.cuda.synchronize()
    prims = ctx.read_gbuffer_prims()
    res = {w: ctx.read_reservoirs(w) for w in (0, 1, 2, 3)}
    fb = ctx.read_framebuffer()
    ctx.close()
    assert np.isfinite(fb).all() and (fb >= 0).all()
    envs = int(((res[3]["flags"] & 1) == 1).sum())
    # the tile with the most hit pixels among a few candidates around the centre
    tile = (W // 2 - TILE // 2, W // 2 + TILE // 2, H // 2 + 20, H // 2 + 20 + TILE)
    numbers = lambda f: [mipt_mod.host_random_number(f * draws + k) for k in range(draws)]  # noqa: E731
    print(f"oracle on tile {tile} ({envs} envmap reservoirs in the frame)", flush=True)
    R = oracle_mod.Renderer(scene5, s, rs, world=world, envmap=env)
    R.render_frames_in_window(frames, tile, numbers)
    sel = np.zeros((H, W), bool)
    sel[tile[2]:tile[3], tile[0]:tile[1]] = True
    sel = sel.reshape(-1)
    assert (prims[sel] >= 0).mean() > 0.5
    assert np.array_equal(prims[sel], R.gb_prims[sel])
    for w in (0, 1, 2, 3):
        a, b = res[w][sel], R.read_reservoirs(w)[sel]
        same = (a.view(np.uint32).reshape(len(a), -1) == b.view(np.uint32).reshape(len(b), -1)).all(axis=1)
        assert same.all(), (w, np.nonzero(~same)[0][:5])
    g, c = fb.reshape(-1, 3)[sel], R.fb.reshape(-1, 3)[sel]
    assert np.array_equal(g.view(np.uint32), c.view(np.uint32))
    assert envs > 0
