"""A moving camera (render_settings.accumulate = 0) on the GPU through the C
ABI, bit for bit against the oracle: four frames of the config-1 scene with
the camera translating and turning between frames (mipt_set_camera before
each mipt_render_sample). Every frame is sample 0 and the reservoirs carry
over, ReSTIR DI temporal reuse reprojecting each pixel's primary hit into
the previous frame's camera (find_temporal_neighbor_index,
UtilsTemporal.h:37-99). Checked every frame: first hits, the four reservoir
buffers, the framebuffer (the frame alone). Variants: temporal + spatial
reuse, the fused pass with light presampling, permutation sampling, no
neighbour search, and the plain path tracer."""
import numpy as np
import pytest

from conftest import look_at

pytestmark = pytest.mark.gpu

W, H = 96, 72
FRAMES = 4


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def cameras(mipt):
    return [look_at(mipt, (0.06 * f, 0.03 * f, 0.98 - 0.05 * f), (0.12 * f, -0.02 * f, -1.0), 60.0, W, H)
            for f in range(FRAMES)]


VARIANTS = {
    "temporal_spatial": dict(),
    "fused_presampling": dict(do_fused_spatiotemporal=1, do_light_presampling=1,
                              light_presampling_number_of_subsets=8, light_presampling_subset_size=16),
    "permutation": dict(use_permutation_sampling=1),
    "no_search": dict(max_neighbor_search_count=0),
    "plain": None,
}


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_moving_camera(mipt_mod, oracle_mod, torch_mod, variant):
    scene = mipt_mod.generate_scene(1, W, H)
    over = VARIANTS[variant]
    lss = mipt_mod.MIPT_LSS_MIS_LIGHT_BSDF if over is None else mipt_mod.MIPT_LSS_RESTIR_DI
    s = mipt_mod.default_render_settings(width=W, height=H, accumulate=0, direct_light_sampling_strategy=lss)
    rs = mipt_mod.default_restir_di_settings(**(over or {}))
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    if over is not None:
        ctx.set_restir_di_settings(rs)
    ctx.set_render_settings(s)
    R = oracle_mod.Renderer(scene, s, rs if over is not None else None)
    draws = None
    for f, cam in enumerate(cameras(mipt_mod)):
        ctx.set_camera(cam)
        if draws is None:
            draws = ctx.draws_per_sample()
        print(f"{variant}: frame {f}", flush=True)
        ctx.render_sample(f)
        torch_mod.cuda.synchronize()
        R.set_camera(cam)
        R.render_frame(f, [mipt_mod.host_random_number(f * draws + k) for k in range(draws)])
        assert np.array_equal(ctx.read_gbuffer_prims(), R.gb_prims), f
        if over is not None:
            for which in (0, 1, 2, 3):
                g, c = ctx.read_reservoirs(which), R.read_reservoirs(which)
                same = (g.view(np.uint32).reshape(len(g), -1) == c.view(np.uint32).reshape(len(c), -1)).all(axis=1)
                assert same.all(), (f, which, np.nonzero(~same)[0][:5])
        fb = ctx.read_framebuffer()
        assert np.array_equal(fb.view(np.uint32), R.fb.view(np.uint32)), f
    ctx.close()
    assert fb.mean() > 0.0


def test_static_camera_not_accumulating(mipt_mod, torch_mod):
    """Not accumulating, each render_sample call overwrites the framebuffer
    with one frame (sample 0), whatever sample number it passes."""
    scene = mipt_mod.generate_scene(1, W, H)
    s = mipt_mod.default_render_settings(width=W, height=H, accumulate=0)
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    ctx.set_render_settings(s)
    ctx.set_camera(scene.camera)
    sums = []
    for f in range(3):
        ctx.render_sample(f)
        torch_mod.cuda.synchronize()
        sums.append(ctx.read_framebuffer().astype(np.float64).sum())
    ctx.close()
    # one sample's worth each time, not a growing sum
    assert max(sums) < 1.5 * min(sums)
