"""Scenes read from glTF files (mipt_scene_load_gltf) rendered on the GPU
through the C ABI, bit for bit against the oracle on the same loaded arrays:
the reference's PBR Cornell box (glass sphere, metal box, emissive strength
100, smooth vertex normals) and its nested-dielectrics scene, with NEE + MIS
and, on the Cornell box, ReSTIR DI (2 frames). The loader's output itself is
pinned on the host (test_gltf.py)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SCENES = {"cornell_pbr": os.path.join(HERE, "golden", "gltf", "cornell_pbr.gltf"),
          "nested_dielectrics": os.path.join(HERE, "golden", "gltf", "nested-dielectrics.gltf")}
W, H = 160, 90


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


@pytest.mark.parametrize("name", list(SCENES))
def test_gltf_scene_renders_as_the_oracle(mipt_mod, oracle_mod, torch_mod, name):
    scene = mipt_mod.load_gltf(SCENES[name], W, H)
    s = mipt_mod.default_render_settings(width=W, height=H, nb_bounces=6)
    spp = 2
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    ctx.set_render_settings(s)
    ctx.set_camera(scene.camera)
    for sn in range(spp):
        ctx.render_sample(sn)
    fb, fb64 = ctx.read_framebuffer(fp64=True)
    prims = ctx.read_gbuffer_prims()
    ctx.close()
    R = oracle_mod.Renderer(scene, s)
    for sn in range(spp):
        R.render_sample(sn, mipt_mod.host_random_number(2 * sn), mipt_mod.host_random_number(2 * sn + 1))
    print(f"\n{name}: {scene.triangle_count} triangles, {scene.emissive.size} emitters, "
          f"{(prims >= 0).mean() * 100:.1f} % of pixels hit, mean radiance {fb.mean() / spp:.4f}")
    assert np.array_equal(prims, R.gb_prims)
    assert np.array_equal(fb.view(np.uint32), R.fb.view(np.uint32))
    assert np.array_equal(fb64, R.fb64)
    assert fb.mean() > 0.0


def test_gltf_cornell_restir(mipt_mod, oracle_mod, torch_mod):
    scene = mipt_mod.load_gltf(SCENES["cornell_pbr"], W, H)
    s = mipt_mod.default_render_settings(width=W, height=H, nb_bounces=4,
                                         direct_light_sampling_strategy=mipt_mod.MIPT_LSS_RESTIR_DI)
    rs = mipt_mod.default_restir_di_settings()
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    ctx.set_restir_di_settings(rs)
    ctx.set_render_settings(s)
    ctx.set_camera(scene.camera)
    R = oracle_mod.Renderer(scene, s, rs)
    draws = ctx.draws_per_sample()
    for f in range(2):
        ctx.render_sample(f)
        R.render_frame(f, [mipt_mod.host_random_number(f * draws + k) for k in range(draws)])
    for which in (0, 1, 2, 3):
        g, c = ctx.read_reservoirs(which), R.read_reservoirs(which)
        assert np.array_equal(g.view(np.uint32), c.view(np.uint32)), which
    fb = ctx.read_framebuffer()
    ctx.close()
    assert np.array_equal(fb.view(np.uint32), R.fb.view(np.uint32))
