"""The top-down binned-SAH builder (MIPT_BVH_SAH_BINNED) through the C ABI.

Pinned by what a BVH must satisfy whatever its shape:
- structure: every interior node reached once, every triangle in exactly
  one leaf (checked by the library on every build, and here on the copy);
- conservative boxes: each child box holds every vertex of every triangle
  below it (the exact triangle test then never misses a hit);
- quality: SAH cost (C_node = C_tri = 1) at most the LBVH's, fewer node
  fetches per incoherent ray;
- hits: bit-exact with the oracle's octree and brute-force loop, and
  identical to the LBVH's on the same rays;
- determinism: two builds give the same node and triangle arrays.
Edge cases: 1, 2 and 17 triangles, duplicated triangles (every centroid
equal: median splits), a scene above one chunk of 2048 triangles per task
(the multi-block path) and the 1,040,384-triangle scene.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _ctx(mipt_mod, scene, builder):
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh(builder)
    return ctx


def _check_tree(scene, nodes, tris):
    """Walks the copied tree: each triangle in one leaf, each child box
    conservative for the triangles below it. Returns the leaf sizes."""
    refs = nodes[:, 12:14].copy().view(np.int32)
    T = scene.triangle_count
    prims = tris[:, 3].copy().view(np.int32)
    assert sorted(prims.tolist()) == list(range(T))
    V = scene.vertices[scene.indices[prims]]  # (T, 3, 3) in leaf order
    tlo, thi = V.min(axis=1), V.max(axis=1)
    seen = np.zeros(T, np.int32)
    leaf_sizes = []
    # subtree triangle ranges are contiguous in leaf order: compute them
    # bottom-up (children follow their parent in the depth-first layout)
    rng_lo = np.full(len(nodes), T, np.int64)
    rng_hi = np.full(len(nodes), -1, np.int64)
    child_ranges = {}
    for i in range(len(nodes) - 1, -1, -1):
        for c in range(2 if T > 1 else 1):
            r = int(refs[i, c])
            if r < 0:
                first, cnt = (~r) >> 4, ((~r) & 15) + 1
                lo, hi = first, first + cnt - 1
                seen[first:first + cnt] += 1
                leaf_sizes.append(cnt)
            else:
                assert r > i, "a child precedes its parent in the depth-first layout"
                lo, hi = rng_lo[r], rng_hi[r]
            child_ranges[(i, c)] = (lo, hi)
            rng_lo[i] = min(rng_lo[i], lo)
            rng_hi[i] = max(rng_hi[i], hi)
    assert (seen == 1).all()
    for (i, c), (lo, hi) in child_ranges.items():
        n = nodes[i]
        blo = n[0:3] if c == 0 else np.array([n[6], n[7], n[8]])
        bhi = np.array([n[3], n[4], n[5]]) if c == 0 else n[9:12]
        assert (tlo[lo:hi + 1].min(axis=0) >= blo).all() and (thi[lo:hi + 1].max(axis=0) <= bhi).all(), (i, c)
    return leaf_sizes


def _random_rays(mipt_mod, n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, mipt_mod.RAY_DTYPE)
    rays["o"] = rng.uniform(lo, hi, size=(n, 3))
    d = rng.normal(size=(n, 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["last_prim"] = -1
    rays["tmax"] = np.inf
    return rays


def _assert_hits_equal(got, want):
    bad = np.nonzero(got["prim"] != want["prim"])[0]
    assert bad.size == 0, f"{bad.size} hit IDs differ, first at ray {bad[:5]}"
    for f in ("t", "u", "v"):
        assert np.array_equal(np.ascontiguousarray(got[f]).view(np.uint32),
                              np.ascontiguousarray(want[f]).view(np.uint32)), f


def test_sah_tree_config1(mipt_mod, oracle_mod, torch_mod, scene1):
    ctx = _ctx(mipt_mod, scene1, mipt_mod.MIPT_BVH_SAH_BINNED)
    info = ctx.bvh_info()
    assert info.builder == mipt_mod.MIPT_BVH_SAH_BINNED
    assert info.triangle_count == scene1.triangle_count
    assert info.node_count == info.leaf_count - 1
    assert 1 <= info.max_leaf_size <= 16
    nodes, tris = ctx.read_bvh()
    sizes = _check_tree(scene1, nodes, tris)
    assert len(sizes) == info.leaf_count
    # the same tree on a second build
    ctx.build_bvh(mipt_mod.MIPT_BVH_SAH_BINNED)
    n2, t2 = ctx.read_bvh()
    assert n2.tobytes() == nodes.tobytes() and t2.tobytes() == tris.tobytes()
    # hits against the octree and brute force
    rays = _random_rays(mipt_mod, 20000, 1, -1.0, 1.0)
    got = ctx.trace_closest_host(rays)
    _assert_hits_equal(got, oracle_mod.Octree(scene1.vertices, scene1.indices).trace(rays))
    _assert_hits_equal(got[::7], oracle_mod.trace_brute(scene1.vertices, scene1.indices, rays[::7]))
    occ = ctx.trace_any_host(rays)
    assert np.array_equal(occ.astype(bool), got["prim"] >= 0)
    sah = info.sah_cost
    ctx.close()
    lb = _ctx(mipt_mod, scene1, mipt_mod.MIPT_BVH_LBVH)
    print(f"\nconfig 1: SAH cost binned {sah:.2f} vs LBVH {lb.bvh_info().sah_cost:.2f}")
    assert sah <= lb.bvh_info().sah_cost
    lb.close()


def test_sah_full_frame_config2(mipt_mod, oracle_mod, torch_mod, scene2):
    """The benchmark rays (1920x1080 incoherent bounce rays, 1,040,384
    triangles): every hit equal to the LBVH's and to the octree's; fewer
    node fetches per ray and a lower SAH cost than the LBVH."""
    w, h = 1920, 1080
    n = w * h
    out = {}
    for builder in (mipt_mod.MIPT_BVH_LBVH, mipt_mod.MIPT_BVH_SAH_BINNED):
        ctx = _ctx(mipt_mod, scene2, builder)
        d_rays = mipt_mod.empty_device(32 * n)
        d_rng = mipt_mod.empty_device(4 * n)
        d_hits = mipt_mod.empty_device(16 * n)
        d_bounce = mipt_mod.empty_device(32 * n)
        d_counts = mipt_mod.empty_device(8 * n)
        ctx.camera_rays(scene2.camera, w, h, 0, mipt_mod.host_random_number(0), 0, d_rays, d_rng)
        ctx.trace_closest(d_rays, d_hits, n)
        ctx.bounce_rays(d_rays, d_hits, d_rng, d_bounce, n)
        ctx.trace_closest_counted(d_bounce, d_hits, d_counts, n)
        torch_mod.cuda.synchronize()
        counts = mipt_mod.from_device(d_counts, np.uint32)[: 2 * n].reshape(-1, 2).astype(np.float64)
        out[builder] = (mipt_mod.from_device(d_bounce, mipt_mod.RAY_DTYPE)[:n],
                        mipt_mod.from_device(d_hits, mipt_mod.HIT_DTYPE)[:n], counts.mean(axis=0), ctx.bvh_info())
        ctx.close()
    b_l, h_l, c_l, i_l = out[mipt_mod.MIPT_BVH_LBVH]
    b_s, h_s, c_s, i_s = out[mipt_mod.MIPT_BVH_SAH_BINNED]
    assert b_l.tobytes() == b_s.tobytes()  # the same rays (primary hits are equal)
    _assert_hits_equal(h_s, h_l)
    _assert_hits_equal(h_s, oracle_mod.Octree(scene2.vertices, scene2.indices).trace(b_s))
    print(f"\nconfig 2 R_inc: nodes/ray {c_l[0]:.2f} -> {c_s[0]:.2f}, tris/ray {c_l[1]:.2f} -> {c_s[1]:.2f}; "
          f"SAH cost {i_l.sah_cost:.2f} -> {i_s.sah_cost:.2f}; build {i_l.build_ms:.1f} -> {i_s.build_ms:.1f} ms; "
          f"depth {i_l.max_depth} -> {i_s.max_depth}")
    assert i_s.sah_cost <= i_l.sah_cost
    assert c_s[0] < c_l[0]


@pytest.mark.parametrize("case", ["one", "two", "seventeen", "duplicates", "grid_5000"])
def test_sah_edge_cases(mipt_mod, oracle_mod, torch_mod, case):
    rng = np.random.default_rng(5)
    if case == "duplicates":
        # 40 copies of one triangle: every centroid equal, median splits only
        V = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (40, 1))
        I = np.arange(120, dtype=np.int32).reshape(40, 3)
    elif case == "grid_5000":
        # 5000 small triangles on a jittered grid: the root task spans three
        # chunks of 2048 (the multi-block path)
        c = np.stack(np.meshgrid(np.arange(100), np.arange(50), indexing="ij"), -1).reshape(-1, 2).astype(np.float32)
        c = np.concatenate([c, rng.uniform(0, 1, (5000, 1)).astype(np.float32)], 1)
        V = (c[:, None, :] + rng.uniform(-0.4, 0.4, (5000, 3, 3))).astype(np.float32).reshape(-1, 3)
        I = np.arange(15000, dtype=np.int32).reshape(5000, 3)
    else:
        T = {"one": 1, "two": 2, "seventeen": 17}[case]
        V = rng.uniform(-1, 1, (3 * T, 3)).astype(np.float32)
        I = np.arange(3 * T, dtype=np.int32).reshape(T, 3)
    scene = mipt_mod.Scene(V, I)
    ctx = _ctx(mipt_mod, scene, mipt_mod.MIPT_BVH_SAH_BINNED)
    info = ctx.bvh_info()
    assert info.triangle_count == I.shape[0]
    if I.shape[0] > 1:
        nodes, tris = ctx.read_bvh()
        _check_tree(scene, nodes, tris)
    lo, hi = V.min() - 0.5, V.max() + 0.5
    rays = _random_rays(mipt_mod, 4000, 2, lo, hi)
    got = ctx.trace_closest_host(rays)
    _assert_hits_equal(got, oracle_mod.trace_brute(V, I, rays))
    occ = ctx.trace_any_host(rays)
    assert np.array_equal(occ.astype(bool), got["prim"] >= 0)
    ctx.close()


def test_sah_render_equals_lbvh_render(mipt_mod, torch_mod):
    """A rendered frame does not depend on the tree: config 1 at 128x128,
    4 samples, identical framebuffer bits with either builder."""
    scene = mipt_mod.generate_scene(1, 128, 128)
    fbs = []
    for builder in (mipt_mod.MIPT_BVH_LBVH, mipt_mod.MIPT_BVH_SAH_BINNED):
        ctx = _ctx(mipt_mod, scene, builder)
        ctx.set_render_settings(mipt_mod.default_render_settings(width=128, height=128, nb_bounces=4))
        ctx.set_camera(scene.camera)
        for s in range(4):
            ctx.render_sample(s)
        fbs.append(ctx.read_framebuffer())
        ctx.close()
    assert fbs[0].tobytes() == fbs[1].tobytes()
