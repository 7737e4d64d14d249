"""The learned radiance cache inside the path-trace pass (config-1 scene).

- Bookkeeping: every path alive at the query bounce is either a cache query
  or a training record; about 1 in train_every paths trains; no target is
  dropped. A warm-up frame (mipt_radiance_cache_warmup) runs only the paths
  that train, about 1 in warmup_train_every, answers no query and leaves the
  framebuffer as it was.
- Gain: after warm-up frames, the cache form's relMSE against an
  FP64-accumulated high-spp plain render is below 0.8x the plain form's at
  the same sample count (measured 0.71x: the cache removes the variance of
  bounces 2 and later, not that of the first two vertices). relMSE = mean
  over pixels and channels of (x - r)^2 / (r^2 + 1e-2), reduced in float64.
  The equal-time comparison runs at 1920x1080 in bench.py, where the GPU
  work, not the launch rate of a 256x256 frame, sets the time.
- Cache off leaves the plain passes as they are (their bit-exact parity is
  covered by test_gpu_pathtrace.py; here the plain form is only the yardstick).
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, H = 256, 256


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


@pytest.fixture(scope="module")
def scene_small(mipt_mod):
    return mipt_mod.generate_scene(1, W, H)


def _ctx(mipt_mod, scene, cache):
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    if cache:
        ctx.set_radiance_cache(mipt_mod.default_radiance_cache_settings(batch_size=1 << 14))
    s = mipt_mod.default_render_settings(width=W, height=H, nb_bounces=4, accumulate_fp64=1,
                                         radiance_cache=1 if cache else 0)
    ctx.set_render_settings(s)
    ctx.set_camera(scene.camera)
    return ctx


def _render(ctx, torch, spp, first=0):
    for i in range(spp):
        ctx.render_sample(first + i if i else 0)
    torch.cuda.synchronize()
    _, fb64 = ctx.read_framebuffer(fp64=True)
    return fb64 / spp


def relmse(x, r):
    return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))


@pytest.fixture(scope="module")
def reference(mipt_mod, torch_mod, scene_small):
    ctx = _ctx(mipt_mod, scene_small, cache=False)
    ref = _render(ctx, torch_mod, 1024)
    ctx.close()
    return ref


def test_cache_records(mipt_mod, torch_mod, scene_small):
    ctx = _ctx(mipt_mod, scene_small, cache=True)
    ctx.render_sample(0)
    torch_mod.cuda.synchronize()
    queries, train, dropped, loss = ctx.cache_stats()
    n = W * H
    print(f"queries {queries}, training records {train}, dropped {dropped}, loss {loss:.4g}")
    assert dropped == 0
    assert 0.3 * n < queries + train <= n
    frac = train / (queries + train)
    every = mipt_mod.default_radiance_cache_settings().train_every
    assert abs(frac * every - 1.0) < 0.25
    assert np.isfinite(loss) and loss > 0
    ctx.close()


def test_warmup_frame(mipt_mod, torch_mod, scene_small):
    ctx = _ctx(mipt_mod, scene_small, cache=True)
    ctx.render_sample(0)
    torch_mod.cuda.synchronize()
    before = ctx.read_framebuffer(fp64=True)
    params_before = ctx.cache_params()
    for i in range(3):
        ctx.cache_warmup(100 + i)
    torch_mod.cuda.synchronize()
    queries, train, dropped, loss = ctx.cache_stats()
    after = ctx.read_framebuffer(fp64=True)
    params_after = ctx.cache_params()
    ctx.close()
    n = W * H
    every = mipt_mod.default_radiance_cache_settings().warmup_train_every
    print(f"warm-up frame: queries {queries}, training records {train}, loss {loss:.4g}")
    assert queries == 0 and dropped == 0
    # the training paths that reach the query vertex (most of those started)
    assert 0.3 * n / every < train < 1.25 * n / every
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    assert np.isfinite(loss) and loss > 0
    # the optimiser stepped: the parameters moved
    assert not np.array_equal(params_before[0], params_after[0])
    assert not np.array_equal(params_before[2], params_after[2])


def test_cache_gain_at_equal_spp(mipt_mod, torch_mod, scene_small, reference):
    spp = 16
    plain = _ctx(mipt_mod, scene_small, cache=False)
    t0 = time.perf_counter()
    img_plain = _render(plain, torch_mod, spp)
    t_plain = time.perf_counter() - t0
    cache = _ctx(mipt_mod, scene_small, cache=True)
    for i in range(64):  # warm-up frames, outside the accumulated image
        cache.cache_warmup(10000 + i)
    torch_mod.cuda.synchronize()
    t0 = time.perf_counter()
    img_cache = _render(cache, torch_mod, spp)
    t_cache = time.perf_counter() - t0
    e_plain, e_cache = relmse(img_plain, reference), relmse(img_cache, reference)
    queries, train, dropped, loss = cache.cache_stats()
    print(f"relMSE at {spp} spp: plain {e_plain:.4g} ({t_plain * 1e3:.1f} ms), "
          f"cache {e_cache:.4g} ({t_cache * 1e3:.1f} ms); train loss {loss:.4g}; "
          f"mean radiance plain {img_plain.mean():.4f} cache {img_cache.mean():.4f} ref {reference.mean():.4f}")
    assert np.isfinite(img_cache).all() and (img_cache >= 0).all()
    assert e_cache < 0.8 * e_plain
    # the cache form converges to about the plain estimator's mean (its bias is small)
    assert abs(img_cache.mean() - reference.mean()) < 0.05 * reference.mean()
