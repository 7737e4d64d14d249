"""Batched camera views (configuration 3), graph replay and the traversal
autotuner, through the C ABI.

- A batch of views of uneven size renders, pixel for pixel and bit for bit,
  what each view renders alone: every pixel draws the random numbers of its
  index in its own view, so batching changes only how the work is launched
  (one wavefront over the pixels of every view).
- A captured hipGraph replays plain samples with the result of the eager
  launches: checked after every replay, with eager work and synchronisation
  between replays, for a batch of views and for one view.
- The autotuner times every (block, LDS stack, queue tile) configuration and
  leaves the hits unchanged.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _small_views(mipt_mod, count, scale):
    """The first `count` views of the configuration-3 batch, their sizes
    divided by `scale` (the cameras keep their matrices)."""
    views = mipt_mod.generate_views(count)
    for v in views:
        v.width = max(8, v.width // scale)
        v.height = max(8, v.height // scale)
    return views


def _ctx(mipt_mod, scene, settings):
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    ctx.set_render_settings(settings)
    return ctx


def test_batch_equals_each_view_alone(mipt_mod, torch_mod):
    scene = mipt_mod.generate_scene(1, 64, 64)
    views = _small_views(mipt_mod, 6, 6)
    sizes = [(v.width, v.height) for v in views]
    s = mipt_mod.default_render_settings(nb_bounces=5)
    ctx = _ctx(mipt_mod, scene, s)
    ctx.set_views(views)
    for sn in range(3):
        ctx.render_sample(sn)
    fb, fb64 = ctx.read_framebuffer(fp64=True)
    ctx.close()
    assert fb.size == 3 * sum(w * h for w, h in sizes)
    off = 0
    for v, (w, h) in zip(views, sizes):
        one = _ctx(mipt_mod, scene, mipt_mod.default_render_settings(width=w, height=h, nb_bounces=5))
        one.set_camera(v.camera)
        for sn in range(3):
            one.render_sample(sn)
        f1, f164 = one.read_framebuffer(fp64=True)
        one.close()
        n = 3 * w * h
        assert fb[off:off + n].tobytes() == f1.tobytes(), (w, h)
        assert fb64[off:off + n].tobytes() == f164.tobytes(), (w, h)
        off += n
    assert fb.mean() > 0.01


def test_views_refuse_restir_and_cache(mipt_mod, torch_mod):
    scene = mipt_mod.generate_scene(1, 64, 64)
    views = _small_views(mipt_mod, 2, 8)
    ctx = _ctx(mipt_mod, scene, mipt_mod.default_render_settings(width=64, height=64,
                                                                  direct_light_sampling_strategy=5))
    with pytest.raises(mipt_mod.MiptError):
        ctx.set_views(views)
    ctx.set_render_settings(mipt_mod.default_render_settings(width=64, height=64))
    ctx.set_views(views)
    with pytest.raises(mipt_mod.MiptError):
        ctx.set_render_settings(mipt_mod.default_render_settings(radiance_cache=1))
    ctx.set_views([])
    ctx.set_render_settings(mipt_mod.default_render_settings(width=64, height=64))
    ctx.set_camera(scene.camera)
    ctx.render_sample(0)
    assert ctx.read_framebuffer().size == 3 * 64 * 64
    ctx.close()


@pytest.mark.parametrize("batch", [True, False])
def test_graph_replays_equal_eager(mipt_mod, torch_mod, batch):
    scene = mipt_mod.generate_scene(1, 96, 64)
    s = mipt_mod.default_render_settings(width=96, height=64, nb_bounces=5)
    views = _small_views(mipt_mod, 5, 8)
    graph = _ctx(mipt_mod, scene, s)
    eager = _ctx(mipt_mod, scene, s)
    for c in (graph, eager):
        if batch:
            c.set_views(views)
        else:
            c.set_camera(scene.camera)
    graph.set_graph_capture(True)
    for sn in range(6):
        graph.render_sample(sn)
        eager.render_sample(sn)
        torch_mod.cuda.synchronize()
        a, a64 = graph.read_framebuffer(fp64=True)
        b, b64 = eager.read_framebuffer(fp64=True)
        assert a.tobytes() == b.tobytes() and a64.tobytes() == b64.tobytes(), f"sample {sn}"
        assert graph.graph_replays() == sn  # sample 0 eager, then one replay per sample
    # a change of settings is recaptured before the next replay
    s2 = mipt_mod.default_render_settings(width=96, height=64, nb_bounces=2)
    for c in (graph, eager):
        c.set_render_settings(s2)
        if not batch:
            c.set_camera(scene.camera)
    for sn in range(3):
        graph.render_sample(sn)
        eager.render_sample(sn)
        torch_mod.cuda.synchronize()
        assert graph.read_framebuffer().tobytes() == eager.read_framebuffer().tobytes(), f"after change, {sn}"
    assert graph.graph_replays() == 2
    graph.close()
    eager.close()


def test_autotune_trace(mipt_mod, torch_mod, scene1):
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene1)
    ctx.build_bvh()
    w = h = 256
    n = w * h
    d_cam = mipt_mod.empty_device(32 * n)
    d_rng = mipt_mod.empty_device(4 * n)
    d_hits = mipt_mod.empty_device(16 * n)
    d_rays = mipt_mod.empty_device(32 * n)
    ctx.camera_rays(scene1.camera, w, h, 0, mipt_mod.host_random_number(0), 0, d_cam, d_rng)
    ctx.trace_closest(d_cam, d_hits, n)
    ctx.bounce_rays(d_cam, d_hits, d_rng, d_rays, n)
    ctx.trace_closest(d_rays, d_hits, n)
    torch_mod.cuda.synchronize()
    before = mipt_mod.from_device(d_hits, mipt_mod.HIT_DTYPE)[:n].copy()
    best, table = ctx.autotune_trace(d_rays, n, repeats=2)
    assert best.configurations == 36
    assert np.isfinite(table).all() and (table > 0).all()
    assert best.best_ms == pytest.approx(table.min())
    assert best.block_size in (64, 128, 256) and best.lds_stack in (8, 16, 24, 32) and best.queue_tile in (1, 2, 4)
    print(f"\nautotune: block {best.block_size}, stack {best.lds_stack}, tile {best.queue_tile}: "
          f"{best.best_ms * 1e3:.1f} us (default {best.default_ms * 1e3:.1f} us)")
    ctx.trace_closest(d_rays, d_hits, n)
    torch_mod.cuda.synchronize()
    after = mipt_mod.from_device(d_hits, mipt_mod.HIT_DTYPE)[:n]
    assert after.tobytes() == before.tobytes()
    for tile in (2, 4):
        ctx.set_queue_tile(tile)
        ctx.trace_closest(d_rays, d_hits, n)
        torch_mod.cuda.synchronize()
        assert mipt_mod.from_device(d_hits, mipt_mod.HIT_DTYPE)[:n].tobytes() == before.tobytes()
    with pytest.raises(mipt_mod.MiptError):
        ctx.set_queue_tile(3)
    ctx.close()
