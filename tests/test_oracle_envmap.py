"""The oracle's ambient light (WorldSettings, Envmap.h, Packing.h:441-485),
pinned by what the format and the mathematics fix:

- R9G9B9E5 packing against an independent implementation of the published
  shared-exponent format (EXT_texture_shared_exponent: 9-bit mantissas,
  5-bit exponent, bias 15), with the reference's round-half-even on ties;
- the envmap's alias table reproducing the texels' luminance pmf, and its
  sum being the row-order binary32 sum;
- sampling and evaluation agreeing on the texel, radiance and pdf of every
  sampled direction under a rotation of the map;
- closed forms: a Lambertian floor under a sky of constant radiance L_up
  above the horizon reflects albedo x L_up (MIS of the envmap light sample
  and the BSDF sample), and under a uniform ambient colour c albedo x c
  after one bounce; camera rays that miss see the map unscaled by its
  intensity (or scaled, with envmap_scale_background_intensity).
"""
import numpy as np
import pytest

from conftest import look_at, sky_scene

W = H = 32


def spec_pack(rgb):
    """The published R9G9B9E5 encoding in float64 (N = 9, B = 15, Emax = 31),
    rounding half up; returns (q_r, q_g, q_b, E)."""
    rgb = np.clip(np.asarray(rgb, np.float64), 0.0, 511.0 / 512.0 * 2.0 ** 16)
    m = rgb.max()
    e = max(-16, int(np.floor(np.log2(m)))) if m > 0 else -16
    E = e + 16
    if np.floor(m / 2.0 ** (E - 24) + 0.5) == 512:
        E += 1
    q = np.floor(rgb / 2.0 ** (E - 24) + 0.5).astype(np.int64)
    return int(q[0]), int(q[1]), int(q[2]), E


def fields(p):
    return p & 0x1FF, (p >> 9) & 0x1FF, (p >> 18) & 0x1FF, p >> 27


def test_rgb9e5_matches_the_shared_exponent_format(oracle_mod):
    rng = np.random.default_rng(3)
    n = 20000
    mag = np.exp2(rng.uniform(-20, 16.5, n))
    vals = mag[:, None] * rng.uniform(0, 1, (n, 3)) ** rng.integers(1, 4, (n, 1))
    vals = np.concatenate([vals, [[0, 0, 0], [65408, 0, 1], [1e9, 2, 3], [1.0, 1.0, 1.0], [2.0 ** -16, 0, 0],
                                  [2.0 ** -24, 2.0 ** -25, 0], [-1.0, 0.5, 0.25]]]).astype(np.float32)
    for v in vals:
        got = fields(oracle_mod.rgb9e5_pack(*v))
        want = spec_pack(v.astype(np.float64))
        if got != want:
            # only an exact tie may differ: half up against half even
            E = got[3]
            units = v.astype(np.float64) / 2.0 ** (E - 24)
            assert np.any(np.abs(units - np.floor(units) - 0.5) == 0), (v, got, want)
        back = oracle_mod.rgb9e5_unpack(oracle_mod.rgb9e5_pack(*v))
        c = np.clip(v.astype(np.float64), 0, 65408)
        unit = 2.0 ** (got[3] - 24)
        assert np.all(np.abs(back - c) <= unit / 2), (v, back)


def test_rgb9e5_ties_round_to_even_and_round_trip(oracle_mod):
    # channel values of 256.5 and 257.5 units of 2^-8 next to a max channel of 1
    u = 2.0 ** -8
    q = fields(oracle_mod.rgb9e5_pack(1.0, 0.5 * u + 0.5 - 0.5 * u * 0, 1.5 * u))
    assert q[3] == 16 and q[0] == 256
    assert fields(oracle_mod.rgb9e5_pack(1.0, 128.5 * u, 129.5 * u))[1:3] == (128, 130)
    # 511.5 units: rounds up to 512, the exponent goes up by one
    p = fields(oracle_mod.rgb9e5_pack(511.5 * u, 0, 0))
    assert p == (256, 0, 0, 17)
    rng = np.random.default_rng(4)
    for _ in range(2000):
        qr, qg, qb = (int(x) for x in rng.integers(0, 512, 3))
        qr = max(qr, 256)  # canonical: the largest channel has its top bit set
        E = int(rng.integers(0, 32))
        p = qr | (qg << 9) | (qb << 18) | (E << 27)
        assert oracle_mod.rgb9e5_pack(*oracle_mod.rgb9e5_unpack(p)) == p


def two_tone(w, h, up, down):
    img = np.zeros((h, w, 3), np.float32)
    img[h // 2:] = up
    img[: h // 2] = down
    return img


def test_envmap_tables(mipt_mod, oracle_mod):
    rng = np.random.default_rng(5)
    env = np.exp(rng.normal(0, 1, (16, 32, 3))).astype(np.float32)
    R = oracle_mod.Renderer(sky_scene(mipt_mod, 8, 8), mipt_mod.default_render_settings(width=8, height=8),
                            envmap=env)
    probas, alias, total, texels = R.env_tables()
    lum = (env[..., 0] * np.float32(0.3086) + env[..., 1] * np.float32(0.6094)) + env[..., 2] * np.float32(0.0820)
    lum = lum.reshape(-1).astype(np.float32)
    s = np.float32(0)
    for x in lum:
        s = np.float32(s + x)
    assert total == s
    pmf = probas.astype(np.float64).copy()
    np.add.at(pmf, alias, 1.0 - probas.astype(np.float64))
    pmf /= lum.size
    # the entries left at the end absorb the rounding of the binary32 sum
    want = lum.astype(np.float64) / lum.astype(np.float64).sum()
    assert np.abs(pmf - want).max() < 2e-6 * max(1.0, want.max() * lum.size) / lum.size * 8
    want = [oracle_mod.rgb9e5_pack(*t) for t in env.reshape(-1, 3)]
    assert texels.tolist() == want


def rotation_y(deg):
    """[column][row] of a rotation by `deg` about +Y, and its inverse."""
    a = np.radians(deg)
    m = np.array([[np.cos(a), 0, -np.sin(a)], [0, 1, 0], [np.sin(a), 0, np.cos(a)]], np.float32)  # rows
    return m.T.copy(), m.copy()


def test_envmap_sample_and_eval_agree(mipt_mod, oracle_mod):
    """The direction of a sampled texel, evaluated, gives back the sample's
    radiance and pdf (the polar angle of the evaluation comes from the world
    direction, the texel from the rotated one) on almost every sample; the
    rare exceptions lie on a texel border."""
    rng = np.random.default_rng(6)
    env = np.exp(rng.normal(0, 1.5, (64, 128, 3))).astype(np.float32)
    to_world, to_env = rotation_y(90.0)
    world = mipt_mod.default_world_settings(ambient_light_type=2, envmap_intensity=1.5,
                                            envmap_to_world=to_world, world_to_envmap=to_env)
    R = oracle_mod.Renderer(sky_scene(mipt_mod, 8, 8), mipt_mod.default_render_settings(width=8, height=8),
                            world=world, envmap=env)
    st = 12345
    same_rad = same_pdf = 0
    n = 4000
    for _ in range(n):
        d, r, pdf, st = R.env_sample(st)
        assert abs(np.linalg.norm(d) - 1) < 1e-5 and pdf > 0
        r2, pdf2 = R.env_eval(d)
        same_rad += np.array_equal(r, r2)
        same_pdf += abs(pdf2 - pdf) <= 1e-4 * pdf
    assert same_rad >= 0.97 * n and same_pdf >= 0.97 * n
    # the same map under the transposed matrix pair: evaluation no longer
    # inverts the sampler's rotation
    world_t = mipt_mod.default_world_settings(ambient_light_type=2, envmap_intensity=1.5,
                                              envmap_to_world=to_world, world_to_envmap=to_world)
    Rt = oracle_mod.Renderer(sky_scene(mipt_mod, 8, 8), mipt_mod.default_render_settings(width=8, height=8),
                             world=world_t, envmap=env)
    st = 12345
    agree = 0
    for _ in range(500):
        d, r, pdf, st = Rt.env_sample(st)
        agree += np.array_equal(Rt.env_eval(d)[0], r)
    assert agree < 0.2 * 500


def floor_scene(mipt_mod, w, h):
    """The 2 x 2 Lambertian floor alone (albedo 0.73), seen from above its
    far edge: the floor fills the lower part of the image, the sky the top."""
    full = sky_scene(mipt_mod, w, h, lights=False)
    cam = look_at(mipt_mod, (0.0, 0.4, 2.0), (0.0, -1.0, 0.0), 70.0, w, h)
    return mipt_mod.Scene(full.vertices, full.indices[:4], full.material_indices[:4], full.materials,
                          np.zeros(0, np.int32), cam, w, h)


def interior(hit):
    """Hit pixels whose 8 neighbours are hits too: under camera jitter the
    floor's border pixels also see the sky in some frames."""
    m = hit.reshape(H, W)
    core = m.copy()
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            core &= np.roll(np.roll(m, dy, 0), dx, 1)
    core[0, :] = core[-1, :] = core[:, 0] = core[:, -1] = False
    return core.reshape(-1)


def render(mipt_mod, oracle_mod, scene, frames, world=None, envmap=None, **settings):
    s = mipt_mod.default_render_settings(width=W, height=H, **settings)
    R = oracle_mod.Renderer(scene, s, world=world, envmap=envmap)
    for f in range(frames):
        R.camera_pass(f, mipt_mod.host_random_number(2 * f))
        R.path_trace(f, mipt_mod.host_random_number(2 * f + 1))
    return R.fb64.reshape(-1, 3) / frames, R


UP = np.float32([1.0, 0.8, 0.5])
DOWN = np.float32([0.05, 0.1, 0.2])


@pytest.mark.parametrize("case", ["plain", "intensity", "flipped"])
def test_floor_under_a_two_tone_sky(mipt_mod, oracle_mod, case):
    scene = floor_scene(mipt_mod, W, H)
    albedo = np.float64(scene.materials[0].base_color[0])
    env = two_tone(64, 32, UP, DOWN)
    kw = {}
    intensity = 1.0
    lit = UP
    if case == "intensity":
        intensity = 2.0
        kw = dict(envmap_intensity=2.0)
    elif case == "flipped":
        # 180 degrees about X: the map's lower half is now the sky
        flip = np.diag([1.0, -1.0, -1.0]).astype(np.float32)
        kw = dict(envmap_to_world=flip, world_to_envmap=flip)
        lit = DOWN
    world = mipt_mod.default_world_settings(ambient_light_type=2, **kw)
    img, R = render(mipt_mod, oracle_mod, scene, 48, world=world, envmap=env, nb_bounces=0)
    floor = R.gb_prims >= 0
    assert 0.2 < floor.mean() < 0.95
    want = albedo * intensity * lit.astype(np.float64)
    core = interior(floor)
    assert core.sum() > 100
    assert np.allclose(img[core].mean(axis=0), want, rtol=0.01), (img[core].mean(axis=0), want)
    # one frame: each background pixel is one of the two tones exactly
    # (packed values), unscaled by the intensity
    img1, R1 = render(mipt_mod, oracle_mod, scene, 1, world=world, envmap=env, nb_bounces=0)
    tones = {tuple(oracle_mod.rgb9e5_unpack(oracle_mod.rgb9e5_pack(*t)) * np.float32(intensity) /
                   np.float32(intensity)) for t in (UP, DOWN)}
    bg = {tuple(np.float32(p)) for p in img1[R1.gb_prims < 0]}
    assert bg <= tones and len(bg) >= 1


def test_background_scaled_by_intensity(mipt_mod, oracle_mod):
    scene = floor_scene(mipt_mod, W, H)
    env = two_tone(64, 32, UP, DOWN)
    world = mipt_mod.default_world_settings(ambient_light_type=2, envmap_intensity=3.0,
                                            envmap_scale_background_intensity=1)
    img, R = render(mipt_mod, oracle_mod, scene, 1, world=world, envmap=env, nb_bounces=0)
    bg = img[R.gb_prims < 0]
    up = oracle_mod.rgb9e5_unpack(oracle_mod.rgb9e5_pack(*UP)) * np.float32(3.0)
    down = oracle_mod.rgb9e5_unpack(oracle_mod.rgb9e5_pack(*DOWN)) * np.float32(3.0)
    assert all(np.array_equal(np.float32(p), up) or np.array_equal(np.float32(p), down) for p in bg)


def test_uniform_ambient_light(mipt_mod, oracle_mod):
    """A uniform ambient colour is not sampled: it reaches the floor through
    the continuation ray that leaves the scene (albedo x c on average) and
    camera rays that miss see it exactly."""
    scene = floor_scene(mipt_mod, W, H)
    albedo = np.float64(scene.materials[0].base_color[0])
    c = np.float32([0.4, 0.6, 0.9])
    world = mipt_mod.default_world_settings(ambient_light_type=1, uniform_light_color=c)
    img, R = render(mipt_mod, oracle_mod, scene, 64, world=world, nb_bounces=1)
    floor = R.gb_prims >= 0
    assert np.allclose(img[interior(floor)].mean(axis=0), albedo * c, rtol=0.01)
    img1, R1 = render(mipt_mod, oracle_mod, scene, 1, world=world, nb_bounces=1)
    miss = R1.gb_prims < 0
    assert np.array_equal(np.float32(img1[miss]), np.broadcast_to(c, img1[miss].shape))
    # without a bounce nothing reaches the floor
    img0, R0 = render(mipt_mod, oracle_mod, scene, 1, world=world, nb_bounces=0)
    assert (img0[R0.gb_prims >= 0] == 0).all()
    # no ambient light at all
    none = mipt_mod.default_world_settings(ambient_light_type=0, uniform_light_color=c)
    img_n, _ = render(mipt_mod, oracle_mod, scene, 1, world=none, nb_bounces=1)
    assert (img_n == 0).all()


def restir_render(mipt_mod, oracle_mod, scene, frames, world, env, strategy=5, w=W, h=H, **restir):
    s = mipt_mod.default_render_settings(width=w, height=h, nb_bounces=0, direct_light_sampling_strategy=strategy)
    rs = mipt_mod.default_restir_di_settings(**restir)
    R = oracle_mod.Renderer(scene, s, rs, world=world, envmap=env)
    if strategy == 5:
        draws = 3 + (3 + rs.spatial_number_of_passes - 1 if rs.do_fused_spatiotemporal else
                     2 * rs.do_temporal_reuse_pass + rs.spatial_number_of_passes * rs.do_spatial_reuse_pass)
    else:
        draws = 2
    for f in range(frames):
        R.render_frame(f, [mipt_mod.host_random_number(f * draws + k) for k in range(draws)])
    return R.fb64.reshape(-1, 3) / frames, R


def test_restir_floor_under_a_two_tone_sky(mipt_mod, oracle_mod):
    """ReSTIR DI initial candidates, all envmap samples (no emissive
    triangle): the floor converges to albedo x L_up as with MIS."""
    scene = floor_scene(mipt_mod, W, H)
    albedo = np.float64(scene.materials[0].base_color[0])
    world = mipt_mod.default_world_settings(ambient_light_type=2)
    env = two_tone(64, 32, UP, DOWN)
    img, R = restir_render(mipt_mod, oracle_mod, scene, 64, world, env, do_temporal_reuse_pass=0,
                           do_spatial_reuse_pass=0)
    floor = R.gb_prims >= 0
    res = R.read_reservoirs(3)
    assert ((res["flags"][floor] & 1) == 1)[res["ucw"][floor] > 0].all()  # envmap samples
    assert np.allclose(img[interior(floor)].mean(axis=0), albedo * UP.astype(np.float64), rtol=0.01)


@pytest.mark.parametrize("case", ["initial", "temporal_spatial", "fused_presampled", "envmap_only"])
def test_restir_with_envmap_unbiased(mipt_mod, oracle_mod, case):
    """ReSTIR DI over emissive triangles and the envmap (candidates drawn from
    either, BSDF candidates that leave the scene become envmap samples)
    averages to the MIS light + BSDF + envmap image at bounce 0 (1.5 %)."""
    rng = np.random.default_rng(9)
    env = (np.exp(rng.normal(0, 0.5, (32, 64, 3))) * 0.5).astype(np.float32)
    scene = sky_scene(mipt_mod, W, H, lights=case != "envmap_only")
    world = mipt_mod.default_world_settings(ambient_light_type=2)
    ref, _ = restir_render(mipt_mod, oracle_mod, scene, 384, world, env, strategy=3)
    kw = {"initial": dict(do_temporal_reuse_pass=0, do_spatial_reuse_pass=0),
          "temporal_spatial": dict(),
          "fused_presampled": dict(do_fused_spatiotemporal=1, do_light_presampling=1,
                                   light_presampling_number_of_subsets=16, light_presampling_subset_size=64),
          "envmap_only": dict(do_temporal_reuse_pass=0, do_spatial_reuse_pass=0)}[case]
    img, R = restir_render(mipt_mod, oracle_mod, scene, 256, world, env, **kw)
    assert np.allclose(img.mean(axis=0), ref.mean(axis=0), rtol=0.015), (img.mean(axis=0), ref.mean(axis=0))
    res = R.read_reservoirs(3)
    live = res["ucw"] > 0
    envs = (res["flags"] & 1) == 1
    assert (res["tri"][live & envs] == -1).all()
    if case == "envmap_only":
        assert envs[live].all()
    else:
        assert envs[live].any() and (~envs[live]).any()
