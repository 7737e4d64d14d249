"""Pins the oracle's Moller-Trumbore test (Triangle.h:20-63), brute-force
traversal and octree (BVH.h, BoundingVolume.h) against known answers, an
independent float64 plane/barycentric intersection and each other."""
import numpy as np
import pytest


def _ray(o, d, last_prim=-1, tmax=np.inf):
    r = np.zeros(1, dtype=[("o", "<f4", (3,)), ("last_prim", "<i4"), ("d", "<f4", (3,)), ("tmax", "<f4")])
    r["o"] = o
    r["d"] = d
    r["last_prim"] = last_prim
    r["tmax"] = tmax
    return r[0]


A, B, Cc = np.array([0, 0, 0], np.float32), np.array([1, 0, 0], np.float32), np.array([0, 1, 0], np.float32)


def test_mt_known_answers(oracle_mod):
    hit, t, u, v = oracle_mod.mt_intersect(A, B, Cc, _ray([0.25, 0.25, 1.0], [0, 0, -1]))
    assert hit and t == 1.0 and u == 0.25 and v == 0.25
    # From below the plane, direction up: still a hit (MT is two-sided).
    hit, t, u, v = oracle_mod.mt_intersect(A, B, Cc, _ray([0.5, 0.25, -2.0], [0, 0, 1]))
    assert hit and t == 2.0 and u == 0.5 and v == 0.25
    # Behind the origin: line hit but t < 0 -> no hit.
    assert not oracle_mod.mt_intersect(A, B, Cc, _ray([0.25, 0.25, 1.0], [0, 0, 1]))[0]
    # Parallel to the plane.
    assert not oracle_mod.mt_intersect(A, B, Cc, _ray([0.25, 0.25, 1.0], [1, 0, 0]))[0]
    # Outside: u + v > 1.
    assert not oracle_mod.mt_intersect(A, B, Cc, _ray([0.75, 0.75, 1.0], [0, 0, -1]))[0]
    # Exactly on the edges u = 0, v = 0 and u + v = 1 counts as a hit.
    assert oracle_mod.mt_intersect(A, B, Cc, _ray([0.0, 0.5, 1.0], [0, 0, -1]))[0]
    assert oracle_mod.mt_intersect(A, B, Cc, _ray([0.5, 0.0, 1.0], [0, 0, -1]))[0]
    assert oracle_mod.mt_intersect(A, B, Cc, _ray([0.5, 0.5, 1.0], [0, 0, -1]))[0]
    # t <= 1e-7 is rejected (origin on the plane).
    assert not oracle_mod.mt_intersect(A, B, Cc, _ray([0.25, 0.25, 0.0], [0, 0, -1]))[0]


def test_mt_determinant_epsilon(oracle_mod):
    # |det| < 1e-7 is rejected even when the ray geometrically crosses the
    # triangle: a 1e-4-sized triangle hit head-on has det = 1e-8.
    s = np.float32(1e-4)
    hit = oracle_mod.mt_intersect(A, B * s, Cc * s, _ray([2.5e-5, 2.5e-5, 1.0], [0, 0, -1]))[0]
    assert not hit
    s = np.float32(1e-3)  # det = 1e-6: accepted
    assert oracle_mod.mt_intersect(A, B * s, Cc * s, _ray([2.5e-4, 2.5e-4, 1.0], [0, 0, -1]))[0]


def _f64_intersect(tri, o, d):
    """Plane intersection + barycentric coordinates in float64."""
    a, b, c = (tri[i].astype(np.float64) for i in range(3))
    n = np.cross(b - a, c - a)
    denom = np.dot(n, d)
    if abs(denom) < 1e-30:
        return None
    t = np.dot(n, a - o) / denom
    p = o + t * d
    # barycentric (u on b, v on c)
    v0, v1, v2 = b - a, c - a, p - a
    d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
    d20, d21 = v2 @ v0, v2 @ v1
    den = d00 * d11 - d01 * d01
    u = (d11 * d20 - d01 * d21) / den
    v = (d00 * d21 - d01 * d20) / den
    return t, u, v


def test_mt_against_float64_geometry(oracle_mod):
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(3000):
        tri = rng.uniform(-1, 1, size=(3, 3)).astype(np.float32)
        o = rng.uniform(-3, 3, size=3).astype(np.float32)
        target = (tri.mean(axis=0) + rng.normal(0, 0.5, size=3)).astype(np.float32)
        d = (target - o).astype(np.float64)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        hit, t, u, v = oracle_mod.mt_intersect(tri[0], tri[1], tri[2], _ray(o, d))
        ref = _f64_intersect(tri, o.astype(np.float64), d.astype(np.float64))
        if ref is None:
            continue
        tr, ur, vr = ref
        margin = min(ur, vr, 1 - ur - vr)
        n = np.cross(tri[1].astype(np.float64) - tri[0], tri[2].astype(np.float64) - tri[0])
        cos = abs(np.dot(n / np.linalg.norm(n), d))
        if abs(margin) < 1e-4 or cos < 1e-2 or abs(tr) < 1e-4:
            continue  # too close to an edge or grazing: rounding decides
        want_hit = margin > 0 and tr > 0
        assert hit == want_hit
        if hit:
            assert abs(t - tr) <= 1e-5 * max(1.0, abs(tr))
            assert abs(u - ur) <= 1e-4 and abs(v - vr) <= 1e-4
        checked += 1
    assert checked > 2000


def _random_scene(rng, ntri, spread=2.0, size=0.3):
    centers = rng.uniform(-spread, spread, size=(ntri, 1, 3))
    verts = (centers + rng.uniform(-size, size, size=(ntri, 3, 3))).astype(np.float32).reshape(-1, 3)
    idx = np.arange(3 * ntri, dtype=np.int32).reshape(-1, 3)
    return verts, idx


def _random_rays(rng, n, spread=3.0):
    rays = np.zeros(n, dtype=[("o", "<f4", (3,)), ("last_prim", "<i4"), ("d", "<f4", (3,)), ("tmax", "<f4")])
    rays["o"] = rng.uniform(-spread, spread, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["last_prim"] = -1
    rays["tmax"] = np.inf
    return rays


def test_brute_force_against_float64(oracle_mod):
    rng = np.random.default_rng(5)
    verts, idx = _random_scene(rng, 60)
    rays = _random_rays(rng, 3000)
    hits = oracle_mod.trace_brute(verts, idx, rays)
    tris = verts[idx]
    agree = 0
    for r, h in zip(rays, hits):
        o, d = r["o"].astype(np.float64), r["d"].astype(np.float64)
        cands = []
        ambiguous = False
        for i in range(len(idx)):
            ref = _f64_intersect(tris[i], o, d)
            if ref is None:
                continue
            t, u, v = ref
            margin = min(u, v, 1 - u - v)
            if abs(margin) < 1e-5 or abs(t) < 1e-5:
                ambiguous = True
            if margin >= 0 and t > 0:
                cands.append((t, i))
        if ambiguous:
            continue
        if not cands:
            assert h["prim"] == -1
        else:
            cands.sort()
            if len(cands) > 1 and cands[1][0] - cands[0][0] < 1e-5:
                continue
            assert h["prim"] == cands[0][1]
            assert abs(h["t"] - cands[0][0]) <= 1e-5 * max(1, cands[0][0])
        agree += 1
    assert agree > 2500


def test_tie_rule_and_filters(oracle_mod):
    # Three copies of one triangle: the closest hit is the lowest index; the
    # self-hit filter skips the ray's own primitive; tmax is exclusive.
    verts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]] * 3, np.float32)
    idx = np.arange(9, dtype=np.int32).reshape(3, 3)
    r = np.array([_ray([0.25, 0.25, 1], [0, 0, -1])])
    assert oracle_mod.trace_brute(verts, idx, r)["prim"][0] == 0
    r["last_prim"] = 0
    assert oracle_mod.trace_brute(verts, idx, r)["prim"][0] == 1
    r["last_prim"] = -1
    r["tmax"] = 1.0  # hit is at exactly t = 1
    assert oracle_mod.trace_brute(verts, idx, r)["prim"][0] == -1
    r["tmax"] = np.nextafter(np.float32(1.0), np.float32(2.0))
    assert oracle_mod.trace_brute(verts, idx, r)["prim"][0] == 0
    assert oracle_mod.trace_brute(verts, idx, r, any_hit=True)["prim"][0] == 0
    tree = oracle_mod.Octree(verts, idx)
    r["tmax"] = np.inf
    assert tree.trace(r)["prim"][0] == 0


def test_empty_scene(oracle_mod):
    verts = np.zeros((0, 3), np.float32)
    idx = np.zeros((0, 3), np.int32)
    rays = _random_rays(np.random.default_rng(1), 10)
    assert (oracle_mod.trace_brute(verts, idx, rays)["prim"] == -1).all()
    assert (oracle_mod.Octree(verts, idx).trace(rays)["prim"] == -1).all()


@pytest.mark.parametrize("seed", [1, 2])
def test_octree_matches_brute_force_random(oracle_mod, seed):
    rng = np.random.default_rng(seed)
    verts, idx = _random_scene(rng, 3000, spread=4.0, size=0.2)
    rays = _random_rays(rng, 4000, spread=5.0)
    tree = oracle_mod.Octree(verts, idx)
    a = tree.trace(rays)
    b = oracle_mod.trace_brute(verts, idx, rays)
    assert np.array_equal(a["prim"], b["prim"])
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))


def test_octree_matches_brute_force_scene1_bounces(oracle_mod, scene1, mipt_mod):
    rays, rng = oracle_mod.camera_rays(scene1.camera, 128, 128, 0, mipt_mod.host_random_number(0))
    tree = oracle_mod.Octree(scene1.vertices, scene1.indices)
    prim = tree.trace(rays)
    bounce, _ = oracle_mod.bounce_rays(scene1.vertices, scene1.indices, rays, prim, rng)
    sub = bounce[::4]
    a = tree.trace(sub)
    b = oracle_mod.trace_brute(scene1.vertices, scene1.indices, sub)
    # The octree's k-DOP slabs are not padded (BoundingVolume.h:106-131), so
    # rounding may in principle cull a grazing hit; count and bound it.
    mismatch = int((a["prim"] != b["prim"]).sum())
    assert mismatch <= 2
    same = a["prim"] == b["prim"]
    assert np.array_equal(a["t"][same].view(np.uint32), b["t"][same].view(np.uint32))


def test_octree_structure(oracle_mod, scene1):
    tree = oracle_mod.Octree(scene1.vertices, scene1.indices)
    st = tree.stats()
    assert st["tri_refs"] == scene1.triangle_count  # each triangle in exactly one leaf
    assert st["max_depth"] <= 32
