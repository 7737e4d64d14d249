"""Pins the oracle's camera rays (HIPRTCamera.h:31-52, CameraRays.h:155-172)
and bounce-ray sampling (Sampling.h:180-201) with float64 restatements of the
geometry and the statistics of a cosine-weighted distribution."""
import numpy as np


def _f64_camera_dir(cam, x, y, w, h):
    iv = np.array(cam.inverse_view, np.float64)
    ip = np.array(cam.inverse_projection, np.float64)

    def xp(m, p):
        q = m @ np.array([p[0], p[1], p[2], 1.0])
        return q[:3] / q[3] if abs(q[3]) >= 1e-10 else q[:3]

    xn, yn = x / w * 2 - 1, y / h * 2 - 1
    o = xp(iv, (0, 0, 0))
    p = xp(iv, xp(ip, (xn, yn, -1.0)))
    d = p - o
    return o, d / np.linalg.norm(d)


def test_camera_ray_matches_float64(oracle_mod, scene1, scene2):
    for scene in (scene1, scene2):
        w, h = scene.width, scene.height
        rng = np.random.default_rng(0)
        for _ in range(200):
            x, y = rng.uniform(0, w), rng.uniform(0, h)
            r = oracle_mod.camera_ray(scene.camera, np.float32(x), np.float32(y), w, h)
            o, d = _f64_camera_dir(scene.camera, np.float32(x), np.float32(y), w, h)
            assert np.allclose(r["o"], o, atol=1e-5)
            assert np.allclose(r["d"], d, atol=2e-6)
            assert abs(np.linalg.norm(r["d"].astype(np.float64)) - 1) < 1e-6


def test_camera_orientation_and_analytic_hit(oracle_mod, scene1):
    # Scene 1 camera at (0, 0, 0.98) looking at -z: the centre ray hits the
    # back wall z = -1 at t = 1.98; y = 0 is the bottom row (d.y < 0).
    cam = scene1.camera
    r = oracle_mod.camera_ray(cam, 256.0, 256.0, 512, 512)
    assert np.allclose(r["d"], [0, 0, -1], atol=1e-6)
    h = oracle_mod.trace_brute(scene1.vertices, scene1.indices, np.array([r]))[0]
    tri = scene1.vertices[scene1.indices[h["prim"]]]
    assert np.allclose(tri[:, 2], -1.0) or h["t"] < 1.98  # back wall, unless a sphere is in front
    bottom = oracle_mod.camera_ray(cam, 256.0, 0.5, 512, 512)
    top = oracle_mod.camera_ray(cam, 256.0, 511.5, 512, 512)
    assert bottom["d"][1] < 0 < top["d"][1]


def test_camera_jitter_draws(oracle_mod, scene1, mipt_mod):
    rn = mipt_mod.host_random_number(0)
    rays, rng = oracle_mod.camera_rays(scene1.camera, 16, 16, 3, rn, 0)
    for pix in (0, 17, 255):
        seed = oracle_mod.pixel_seed(pix, 3, rn, 0)
        draws = oracle_mod.rand01_seq(seed, 2)
        x, y = pix % 16, pix // 16
        want = oracle_mod.camera_ray(scene1.camera, np.float32(x + 0.5) + (draws[0] - np.float32(0.5)),
                                     np.float32(y + 0.5) + (draws[1] - np.float32(0.5)), 16, 16)
        assert rays[pix].tobytes() == want.tobytes()
        assert rng[pix] == oracle_mod.xorshift_seq(seed, 2)[-1]


def test_bounce_rays_cosine_distribution(oracle_mod, scene1, mipt_mod):
    rays, rng = oracle_mod.camera_rays(scene1.camera, 256, 256, 0, mipt_mod.host_random_number(0))
    hits = oracle_mod.Octree(scene1.vertices, scene1.indices).trace(rays)
    bounce, rng2 = oracle_mod.bounce_rays(scene1.vertices, scene1.indices, rays, hits, rng)
    assert (hits["prim"] >= 0).all()
    assert (bounce["last_prim"] == hits["prim"]).all()
    cos = np.empty(len(rays))
    for i in range(len(rays)):
        p, gn, sn = oracle_mod.hit_attributes(scene1.vertices, scene1.indices, rays[i], hits[i])
        if i % 97 == 0:
            assert np.allclose(bounce[i]["o"], p)
        cos[i] = np.dot(bounce[i]["d"].astype(np.float64), sn)
    # Cosine-weighted about the shading normal: cos > 0, E[cos] = 2/3 and
    # cos^2 ~ U[0, 1].
    assert (cos > 0).all()
    assert abs(cos.mean() - 2.0 / 3.0) < 0.005
    c2 = np.sort(cos**2)
    ks = np.max(np.abs(c2 - (np.arange(1, len(c2) + 1) / len(c2))))
    assert ks < 0.01
    assert (rng2 != rng).all()  # two draws consumed


def test_backfacing_normals_face_the_viewer(oracle_mod, scene1, mipt_mod):
    rays, _ = oracle_mod.camera_rays(scene1.camera, 64, 64, 0, mipt_mod.host_random_number(0))
    hits = oracle_mod.trace_brute(scene1.vertices, scene1.indices, rays)
    for i in range(0, len(rays), 13):
        p, gn, sn = oracle_mod.hit_attributes(scene1.vertices, scene1.indices, rays[i], hits[i])
        view = -rays[i]["d"].astype(np.float64)
        assert np.dot(view, gn) >= 0
        assert np.dot(view, sn) >= 0
