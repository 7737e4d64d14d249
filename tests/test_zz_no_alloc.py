"""mipt_trace_closest_noalloc (closest hit with the work counter and spill
stack in the library's static device scratch) against mipt_trace_closest,
through the C ABI: hit records bit-identical on two calls in a row, the
second over fewer rays, on the 1,040,384-triangle scene's 1920x1080
incoherent bounce rays; the same on a context whose first traversal is the
no-allocation one; a BVH deep enough that the grid's spill columns exceed
the scratch (LDS stack of 8 entries) still gives the same hits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _rays(mipt_mod, scene, ctx, w, h):
    n = w * h
    d_cam = mipt_mod.empty_device(32 * n)
    d_rng = mipt_mod.empty_device(4 * n)
    d_hits = mipt_mod.empty_device(16 * n)
    d_rays = mipt_mod.empty_device(32 * n)
    ctx.camera_rays(scene.camera, w, h, 0, mipt_mod.host_random_number(0), 0, d_cam, d_rng)
    ctx.trace_closest(d_cam, d_hits, n)
    ctx.bounce_rays(d_cam, d_hits, d_rng, d_rays, n)
    return d_rays, n


def _hits(mipt_mod, torch_mod, fn, d_rays, n):
    d_hits = mipt_mod.empty_device(16 * n)
    fn(d_rays, d_hits, n)
    torch_mod.cuda.synchronize()
    return mipt_mod.from_device(d_hits, mipt_mod.HIT_DTYPE)[:n]


def _same(a, b):
    return a.tobytes() == b.tobytes()


@pytest.mark.parametrize("builder", ["lbvh", "sah"])
def test_noalloc_equals_trace_closest(mipt_mod, torch_mod, scene2, builder):
    b = mipt_mod.MIPT_BVH_SAH_BINNED if builder == "sah" else mipt_mod.MIPT_BVH_LBVH
    w, h = 1920, 1080
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene2)
    ctx.build_bvh(b)
    d_rays, n = _rays(mipt_mod, scene2, ctx, w, h)
    for count in (n, n // 3 + 17):
        print(f"noalloc {builder}: {count} rays", flush=True)
        want = _hits(mipt_mod, torch_mod, ctx.trace_closest, d_rays, count)
        got = _hits(mipt_mod, torch_mod, ctx.trace_closest_noalloc, d_rays, count)
        assert _same(got, want), f"{count} rays: hit records differ"
        assert (got["prim"] >= 0).mean() > 0.3
    # a deep-spill configuration: 8 LDS entries, the rest in the scratch
    ctx.set_trace_config(128, 8)
    for count in (n, 1000):
        print(f"noalloc {builder}, LDS stack 8: {count} rays", flush=True)
        want = _hits(mipt_mod, torch_mod, ctx.trace_closest, d_rays, count)
        got = _hits(mipt_mod, torch_mod, ctx.trace_closest_noalloc, d_rays, count)
        assert _same(got, want), f"stack 8, {count} rays: hit records differ"
    ctx.close()


def test_noalloc_first_traversal_of_a_context(mipt_mod, oracle_mod, torch_mod):
    """No traversal ran on the context before (nothing was allocated for
    the spill stack): on two calls, the hits equal brute force over every
    triangle (the hit-ID truth) on a strided subset and mipt_trace_closest's
    afterwards on every ray."""
    scene = mipt_mod.generate_scene(1, 256, 256)
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    rng = np.random.default_rng(7)
    rays = np.zeros(50000, mipt_mod.RAY_DTYPE)
    rays["o"] = rng.uniform(-0.9, 0.9, size=(rays.size, 3))
    d = rng.normal(size=(rays.size, 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["last_prim"] = -1
    rays["tmax"] = np.inf
    d_rays = mipt_mod.to_device(rays)
    got = [_hits(mipt_mod, torch_mod, ctx.trace_closest_noalloc, d_rays, count) for count in (rays.size, 4097)]
    want = _hits(mipt_mod, torch_mod, ctx.trace_closest, d_rays, rays.size)
    assert _same(got[0], want) and _same(got[1], want[:4097])
    brute = oracle_mod.trace_brute(scene.vertices, scene.indices, rays[::10])
    assert np.array_equal(got[0]["prim"][::10], brute["prim"])
    assert np.array_equal(got[0]["t"][::10].view(np.uint32), brute["t"].view(np.uint32))
    ctx.close()


def test_noalloc_errors(mipt_mod, torch_mod):
    scene = mipt_mod.generate_scene(1, 64, 64)
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    d = mipt_mod.empty_device(64)
    with pytest.raises(mipt_mod.MiptError):
        ctx.trace_closest_noalloc(d, d, 1)  # no BVH yet
    ctx.build_bvh()
    ctx.trace_closest_noalloc(d, d, 0)  # zero rays: nothing to do
    ctx.close()
