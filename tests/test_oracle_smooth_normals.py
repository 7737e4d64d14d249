"""Pins the oracle's smooth shading normal (get_shading_normal,
Intersect.h:114-133; uv_interpolate, Texture.h:248-251) with what the
definition fixes:
- at the barycentrics (u, v) = (0, 0), (1, 0), (0, 1) the shading normal is
  the normal of vertex A, B, C (a mix-up of the vertices or of u and v
  fails);
- inside the triangle it is normalize(n_B u + n_C v + n_A (1 - u - v)),
  compared against float64 (the binary32 result within 4 ulp-scale of the
  direction);
- three equal vertex normals give that normal everywhere (constant
  interpolation), and a triangle whose first vertex has no normal keeps the
  geometric normal, whatever its other vertices have;
- on a finely tessellated sphere with the exact sphere normals at the
  vertices, the interpolated normal at a hit point is within the
  tessellation's angular error of the true sphere normal there."""
import numpy as np


def _tri_case(rng):
    n = rng.normal(size=(3, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return n.astype(np.float32)


def test_vertices_give_their_normals(oracle_mod):
    rng = np.random.default_rng(1)
    idx = np.array([[0, 1, 2]], np.int32)
    gn = np.array([0, 0, 1], np.float32)
    for _ in range(50):
        vn = _tri_case(rng)
        has = np.ones(3, np.uint8)
        for (u, v), k in (((0.0, 0.0), 0), ((1.0, 0.0), 1), ((0.0, 1.0), 2)):
            got = oracle_mod.shading_normal(idx, vn, has, 0, u, v, gn)
            want = vn[k] / np.float32(np.sqrt(np.float32(vn[k] @ vn[k])))
            assert np.allclose(got, want, atol=2e-7), (u, v, k)


def test_interpolation_against_float64(oracle_mod):
    rng = np.random.default_rng(2)
    idx = np.array([[2, 0, 1]], np.int32)  # A = vertex 2, B = vertex 0, C = vertex 1
    gn = np.array([1, 0, 0], np.float32)
    worst = 0.0
    for _ in range(2000):
        vn = _tri_case(rng)
        has = np.ones(3, np.uint8)
        u, v = rng.uniform(size=2)
        if u + v > 1:
            u, v = 1 - u, 1 - v
        u, v = np.float32(u), np.float32(v)
        got = oracle_mod.shading_normal(idx, vn, has, 0, u, v, gn).astype(np.float64)
        a, b, c = vn[2].astype(np.float64), vn[0].astype(np.float64), vn[1].astype(np.float64)
        w = (1.0 - float(u)) - float(v)
        n = b * float(u) + c * float(v) + a * w
        if np.linalg.norm(n) < 1e-3:
            continue  # nearly cancelling normals: the direction is ill-conditioned
        want = n / np.linalg.norm(n)
        worst = max(worst, float(np.abs(got - want).max() * np.linalg.norm(n)))
    assert worst < 4e-6


def test_constant_normals_and_missing_first_vertex(oracle_mod):
    rng = np.random.default_rng(3)
    idx = np.array([[0, 1, 2]], np.int32)
    gn = np.array([0, 1, 0], np.float32)
    n0 = np.array([0.3, 0.8, -0.2], np.float32)
    n0 /= np.float32(np.linalg.norm(n0))
    vn = np.tile(n0, (3, 1))
    for _ in range(100):
        u, v = rng.uniform(0, 0.5, size=2).astype(np.float32)
        got = oracle_mod.shading_normal(idx, vn, np.ones(3, np.uint8), 0, u, v, gn)
        assert np.allclose(got, n0, atol=3e-7)
        # the first vertex decides (get_shading_normal reads vertex x only)
        got = oracle_mod.shading_normal(idx, vn, np.array([0, 1, 1], np.uint8), 0, u, v, gn)
        assert np.array_equal(got, gn)
        got = oracle_mod.shading_normal(idx, vn, np.array([1, 0, 0], np.uint8), 0, u, v, gn)
        assert np.allclose(got, n0, atol=3e-7)


def test_sphere_normals(mipt_mod, oracle_mod):
    """Icosphere of the config-1 scene (subdivision 4, 5120 triangles), vertex
    normals = normalize(vertex - centre): at random hits the interpolated
    normal is within the tessellation's error of the exact normal of the
    sphere through the hit point (much closer than the flat normal)."""
    scene = mipt_mod.generate_scene(1, 64, 64)
    mi = scene.material_indices
    sphere_mat = [k for k in range(len(scene.materials)) if (mi == k).sum() == 5120][0]
    tris = scene.indices[mi == sphere_mat]
    used = np.unique(tris)
    c = scene.vertices[used].astype(np.float64).mean(axis=0)
    vn = np.zeros_like(scene.vertices)
    d = scene.vertices[used].astype(np.float64) - c
    vn[used] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    has = np.zeros(scene.vertices.shape[0], np.uint8)
    has[used] = 1
    rng = np.random.default_rng(4)
    err_smooth, err_flat = [], []
    for t in rng.choice(np.nonzero(mi == sphere_mat)[0], 300, replace=False):
        u, v = rng.uniform(size=2)
        if u + v > 1:
            u, v = 1 - u, 1 - v
        A, B, C_ = scene.vertices[scene.indices[t]].astype(np.float64)
        p = A + u * (B - A) + v * (C_ - A)
        exact = (p - c) / np.linalg.norm(p - c)
        flat = np.cross(B - A, C_ - A)
        flat /= np.linalg.norm(flat)
        sn = oracle_mod.shading_normal(scene.indices, vn, has, int(t), np.float32(u), np.float32(v),
                                       flat.astype(np.float32)).astype(np.float64)
        err_smooth.append(np.degrees(np.arccos(np.clip(sn @ exact, -1, 1))))
        err_flat.append(np.degrees(np.arccos(np.clip(abs(flat @ exact), -1, 1))))
    assert max(err_smooth) < 0.5 and np.mean(err_smooth) < 0.25 * np.mean(err_flat)
