"""GPU parity of the traversal path through the C ABI (libmipt.so):
camera rays, closest hit, any hit and bounce rays against the oracle.

Bar: bit-exact hit primitive IDs and bit-exact t/u/v against the oracle's
brute-force loop (identical input ray bits); camera rays and bounce rays
bit-exact (sin/cos of the cosine sampling are the fixed-sequence versions on
both sides).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _ctx(mipt_mod, scene):
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    return ctx


def _gpu_camera(mipt_mod, torch_mod, ctx, scene, w, h, sample=0, pass_index=0):
    n = w * h
    d_rays = mipt_mod.empty_device(32 * n)
    d_rng = mipt_mod.empty_device(4 * n)
    ctx.camera_rays(scene.camera, w, h, sample, mipt_mod.host_random_number(pass_index), 0, d_rays, d_rng)
    torch_mod.cuda.synchronize()
    return d_rays, d_rng


def _same_bits(a, b, field):
    return np.array_equal(np.ascontiguousarray(a[field]).view(np.uint32),
                          np.ascontiguousarray(b[field]).view(np.uint32))


def _assert_hits_equal(got, want):
    bad = np.nonzero(got["prim"] != want["prim"])[0]
    assert bad.size == 0, f"{bad.size} hit IDs differ, first at ray {bad[:5]}"
    for f in ("t", "u", "v"):
        assert _same_bits(got, want, f), f"hit {f} bits differ"


def test_camera_rays_bit_exact(mipt_mod, oracle_mod, torch_mod, scene1, scene2):
    for scene, (w, h) in ((scene1, (512, 512)), (scene2, (1920, 1080))):
        ctx = _ctx(mipt_mod, scene)
        for sample, pidx in ((0, 0), (7, 3)):
            d_rays, d_rng = _gpu_camera(mipt_mod, torch_mod, ctx, scene, w, h, sample, pidx)
            got = mipt_mod.from_device(d_rays, mipt_mod.RAY_DTYPE)
            rng = mipt_mod.from_device(d_rng, np.uint32)
            want, want_rng = oracle_mod.camera_rays(scene.camera, w, h, sample, mipt_mod.host_random_number(pidx))
            assert got.tobytes() == want.tobytes()
            assert np.array_equal(rng, want_rng)
        ctx.close()


def test_bvh_info(mipt_mod, scene1, scene2):
    for scene in (scene1, scene2):
        ctx = _ctx(mipt_mod, scene)
        info = ctx.bvh_info()
        assert info.triangle_count == scene.triangle_count
        assert info.leaf_count >= scene.triangle_count // 16
        assert info.node_count == info.leaf_count - 1
        assert 1 <= info.max_leaf_size <= 16
        assert info.sah_cost > 0
        ctx.close()


def test_closest_hit_scene1_bit_exact(mipt_mod, oracle_mod, torch_mod, scene1):
    ctx = _ctx(mipt_mod, scene1)
    w = h = 512
    n = w * h
    d_rays, d_rng = _gpu_camera(mipt_mod, torch_mod, ctx, scene1, w, h)
    d_hits = mipt_mod.empty_device(16 * n)
    ctx.trace_closest(d_rays, d_hits, n)
    d_bounce = mipt_mod.empty_device(32 * n)
    ctx.bounce_rays(d_rays, d_hits, d_rng, d_bounce, n)
    d_hits2 = mipt_mod.empty_device(16 * n)
    ctx.trace_closest(d_bounce, d_hits2, n)
    torch_mod.cuda.synchronize()
    rays = mipt_mod.from_device(d_rays, mipt_mod.RAY_DTYPE)
    hits = mipt_mod.from_device(d_hits, mipt_mod.HIT_DTYPE)
    bounce = mipt_mod.from_device(d_bounce, mipt_mod.RAY_DTYPE)
    hits2 = mipt_mod.from_device(d_hits2, mipt_mod.HIT_DTYPE)

    tree = oracle_mod.Octree(scene1.vertices, scene1.indices)
    # Primary rays: every ray against the brute-force truth on a subset, and
    # against the octree for all rays.
    sub = slice(0, n, 37)
    _assert_hits_equal(hits[sub], oracle_mod.trace_brute(scene1.vertices, scene1.indices, rays[sub]))
    _assert_hits_equal(hits, tree.trace(rays))
    # Bounce rays generated on the GPU, traced on the GPU vs the oracle on
    # the identical ray bits.
    _assert_hits_equal(hits2[sub], oracle_mod.trace_brute(scene1.vertices, scene1.indices, bounce[sub]))
    _assert_hits_equal(hits2, tree.trace(bounce))
    # The GPU bounce-ray generator against the oracle's.
    _, rng = oracle_mod.camera_rays(scene1.camera, w, h, 0, mipt_mod.host_random_number(0))
    want_bounce, want_rng = oracle_mod.bounce_rays(scene1.vertices, scene1.indices, rays, hits, rng)
    assert np.array_equal(bounce["last_prim"], want_bounce["last_prim"])
    assert _same_bits(bounce, want_bounce, "o")
    assert _same_bits(bounce, want_bounce, "d")
    assert np.array_equal(mipt_mod.from_device(d_rng, np.uint32), want_rng)
    ctx.close()


def test_any_hit_scene1_shadow_rays(mipt_mod, oracle_mod, torch_mod, scene1):
    ctx = _ctx(mipt_mod, scene1)
    w = h = 256
    n = w * h
    d_rays, d_rng = _gpu_camera(mipt_mod, torch_mod, ctx, scene1, w, h)
    d_hits = mipt_mod.empty_device(16 * n)
    ctx.trace_closest(d_rays, d_hits, n)
    torch_mod.cuda.synchronize()
    rays = mipt_mod.from_device(d_rays, mipt_mod.RAY_DTYPE)
    hits = mipt_mod.from_device(d_hits, mipt_mod.HIT_DTYPE)
    # Shadow rays from each primary hit to a random point on an emissive
    # triangle, t_max = distance - 1e-4 (Intersect.h:285).
    rng = np.random.default_rng(9)
    em = scene1.emissive[rng.integers(0, len(scene1.emissive), n)]
    tri = scene1.vertices[scene1.indices[em]].astype(np.float32)
    u, v = rng.uniform(size=n), rng.uniform(size=n)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    target = (tri[:, 0] + u[:, None] * (tri[:, 1] - tri[:, 0]) + v[:, None] * (tri[:, 2] - tri[:, 0])).astype(np.float32)
    origin = (rays["o"] + hits["t"][:, None] * rays["d"]).astype(np.float32)
    dvec = (target - origin).astype(np.float64)
    dist = np.linalg.norm(dvec, axis=1)
    shadow = np.zeros(n, mipt_mod.RAY_DTYPE)
    shadow["o"] = origin
    shadow["d"] = (dvec / dist[:, None]).astype(np.float32)
    shadow["last_prim"] = hits["prim"]
    shadow["tmax"] = (dist - 1e-4).astype(np.float32)
    occ = ctx.trace_any_host(shadow)
    want = oracle_mod.trace_brute(scene1.vertices, scene1.indices, shadow[::11], any_hit=True)
    assert np.array_equal(occ[::11].astype(bool), want["prim"] >= 0)
    want_all = oracle_mod.Octree(scene1.vertices, scene1.indices).trace(shadow)
    assert np.array_equal(occ.astype(bool), want_all["prim"] >= 0)
    assert 0.05 < occ.mean() < 0.95
    # Closest hit with the same finite tmax: hits only below tmax.
    got = ctx.trace_closest_host(shadow[::11])
    _assert_hits_equal(got, oracle_mod.trace_brute(scene1.vertices, scene1.indices, shadow[::11]))
    ctx.close()


def test_closest_hit_scene2_full_frame(mipt_mod, oracle_mod, torch_mod, scene2):
    """The benchmark workload: 1920x1080 incoherent bounce rays on the
    1,040,384-triangle scene. All 2,073,600 rays are checked against the
    octree oracle, a strided subset against brute force."""
    ctx = _ctx(mipt_mod, scene2)
    w, h = 1920, 1080
    n = w * h
    d_rays, d_rng = _gpu_camera(mipt_mod, torch_mod, ctx, scene2, w, h)
    d_hits = mipt_mod.empty_device(16 * n)
    ctx.trace_closest(d_rays, d_hits, n)
    d_bounce = mipt_mod.empty_device(32 * n)
    ctx.bounce_rays(d_rays, d_hits, d_rng, d_bounce, n)
    ctx.trace_closest(d_bounce, d_hits, n)
    torch_mod.cuda.synchronize()
    bounce = mipt_mod.from_device(d_bounce, mipt_mod.RAY_DTYPE)
    hits = mipt_mod.from_device(d_hits, mipt_mod.HIT_DTYPE)
    tree = oracle_mod.Octree(scene2.vertices, scene2.indices)
    _assert_hits_equal(hits, tree.trace(bounce))
    sub = slice(0, n, 1013)
    _assert_hits_equal(hits[sub], oracle_mod.trace_brute(scene2.vertices, scene2.indices, bounce[sub]))
    assert 0.05 < (hits["prim"] >= 0).mean() < 0.99
    ctx.close()


@pytest.mark.parametrize("block,stack", [(64, 8), (128, 16), (256, 24), (128, 32), (256, 8)])
def test_launch_configs_agree(mipt_mod, oracle_mod, torch_mod, scene2, block, stack):
    """Every (block size, LDS stack depth) gives identical results; stack 8 on
    the 1M-triangle tree forces the global spill path."""
    ctx = _ctx(mipt_mod, scene2)
    ctx.set_trace_config(block, stack)
    rng = np.random.default_rng(block * 100 + stack)
    n = 50000
    rays = np.zeros(n, mipt_mod.RAY_DTYPE)
    rays["o"] = np.stack([rng.uniform(-30, 30, n), rng.uniform(-2, 8, n), rng.uniform(-30, 30, n)], 1)
    d = rng.normal(size=(n, 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["last_prim"] = -1
    rays["tmax"] = np.inf
    got = ctx.trace_closest_host(rays)
    assert ctx.bvh_info().max_depth > 8
    _assert_hits_equal(got, oracle_mod.Octree(scene2.vertices, scene2.indices).trace(rays))
    occ = ctx.trace_any_host(rays)
    assert np.array_equal(occ.astype(bool), got["prim"] >= 0)
    ctx.close()


def test_counted_trace_matches_and_counts(mipt_mod, torch_mod, scene1):
    ctx = _ctx(mipt_mod, scene1)
    n = 128 * 128
    d_rays, _ = _gpu_camera(mipt_mod, torch_mod, ctx, scene1, 128, 128)
    d_h1 = mipt_mod.empty_device(16 * n)
    d_h2 = mipt_mod.empty_device(16 * n)
    d_c = mipt_mod.empty_device(8 * n)
    ctx.trace_closest(d_rays, d_h1, n)
    ctx.trace_closest_counted(d_rays, d_h2, d_c, n)
    torch_mod.cuda.synchronize()
    assert mipt_mod.from_device(d_h1, np.uint8).tobytes() == mipt_mod.from_device(d_h2, np.uint8).tobytes()
    counts = mipt_mod.from_device(d_c, np.uint32).reshape(-1, 2)
    assert (counts[:, 0] >= 1).all() and (counts[:, 1] >= 1).all()
    ctx.close()


def test_edge_cases(mipt_mod, oracle_mod, torch_mod):
    # Empty scene: every ray misses, any-hit all clear.
    empty = mipt_mod.Scene(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32))
    ctx = _ctx(mipt_mod, empty)
    rays = np.zeros(100, mipt_mod.RAY_DTYPE)
    rays["d"] = [0, 0, 1]
    rays["tmax"] = np.inf
    rays["last_prim"] = -1
    assert (ctx.trace_closest_host(rays)["prim"] == -1).all()
    assert (ctx.trace_any_host(rays) == 0).all()
    ctx.close()

    # One triangle; axis-aligned rays (zero direction components), a ray
    # leaving the triangle itself (self-hit filter), tmax = 0 and exactly t.
    tri = mipt_mod.Scene(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.int32))
    ctx = _ctx(mipt_mod, tri)
    r = np.zeros(6, mipt_mod.RAY_DTYPE)
    r["o"] = [[0.25, 0.25, 1], [0.25, 0.25, 1], [0.25, 0.25, 1], [0.0, 0.0, 1], [0.25, 0.25, 1], [2, 2, 1]]
    r["d"] = [[0, 0, -1], [0, 0, -1], [0, 0, -1], [0, 0, -1], [0, 0, -1], [0, 0, -1]]
    r["last_prim"] = [-1, 0, -1, -1, -1, -1]
    r["tmax"] = [np.inf, np.inf, 0.0, np.inf, 1.0, np.inf]
    got = ctx.trace_closest_host(r)
    want = oracle_mod.trace_brute(tri.vertices, tri.indices, r)
    _assert_hits_equal(got, want)
    assert list(got["prim"]) == [0, -1, -1, 0, -1, -1]
    ctx.close()

    # Duplicated triangles: ties in t go to the lowest primitive index.
    dup = mipt_mod.Scene(np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (5, 1)),
                         np.arange(15, dtype=np.int32).reshape(5, 3)[::-1].copy())
    ctx = _ctx(mipt_mod, dup)
    r1 = r[:1].copy()
    got = ctx.trace_closest_host(r1)
    assert got["prim"][0] == 0
    r1["last_prim"] = 0
    assert ctx.trace_closest_host(r1)["prim"][0] == 1
    ctx.close()


def test_zero_rays_and_errors(mipt_mod, torch_mod, scene1):
    ctx = mipt_mod.Context(0)
    d = mipt_mod.empty_device(64)
    with pytest.raises(mipt_mod.MiptError):
        ctx.trace_closest(d, d, 1)  # no BVH yet
    with pytest.raises(mipt_mod.MiptError):
        ctx.build_bvh()  # no scene yet
    bad = mipt_mod.Scene(np.zeros((3, 3), np.float32), np.array([[0, 1, 5]], np.int32))
    with pytest.raises(mipt_mod.MiptError):
        ctx.set_scene(bad)  # vertex index out of range
    ctx.set_scene(scene1)
    ctx.build_bvh()
    ctx.trace_closest(d, d, 0)  # n = 0 is a no-op
    with pytest.raises(mipt_mod.MiptError):
        ctx.set_trace_config(100, 16)
    ctx.close()


def test_deterministic_repeat(mipt_mod, torch_mod, scene2):
    ctx = _ctx(mipt_mod, scene2)
    n = 1920 * 1080
    d_rays, d_rng = _gpu_camera(mipt_mod, torch_mod, ctx, scene2, 1920, 1080)
    d_h = mipt_mod.empty_device(16 * n)
    ctx.trace_closest(d_rays, d_h, n)
    d_b = mipt_mod.empty_device(32 * n)
    ctx.bounce_rays(d_rays, d_h, d_rng, d_b, n)
    outs = []
    for _ in range(3):
        ctx.trace_closest(d_b, d_h, n)
        torch_mod.cuda.synchronize()
        outs.append(mipt_mod.from_device(d_h, np.uint8).tobytes())
    assert outs[0] == outs[1] == outs[2]
    ctx.close()
