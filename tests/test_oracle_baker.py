"""The oracle's restatement of the energy-compensation LUT baker
(src/Device/kernels/Baking/, src/Renderer/Baker/GPUBakerKernel.cpp:52-100),
pinned by the tables the reference baked and ships (data/BRDFsData/, embedded
in the library and RGBE-decoded):

- drawn with the reference's launch plan and seeds (launch i seeds texel t's
  xorshift32 with wang_hash(t + 1) * i) and summed in its order, a texel of
  every renderer table (conductor in both masking flavours, glossy
  dielectric, glass entering and exiting, thin glass), passed through the
  Radiance file's 8-bit truncation, equals the shipped value: 150 random
  texels per table at the reference's sample counts, at least 97 % of them
  exactly and 98 % within one mantissa step (over 1000 texels a table:
  99.8-100 % exact, the rest at grazing cos_theta_o);
- the comparison tells the flavours apart (an uncorrelated bake against the
  correlated table matches about half the texels);
- closed forms: the Fresnel table at F0 = 0 is 0, the conductor's perfectly
  smooth rows integrate 1 (up to the drift of the sequential binary32 sum),
  the glossy dielectric at relative IOR 1 (no Fresnel) is the white
  Lambertian's albedo 1 within Monte Carlo noise, and every albedo of the
  lossless lobes is at most 1 plus that noise;
- the launch plan (iterations per launch, launches, normalising count) of
  each table at its default settings, from the reference's arithmetic.
"""
import numpy as np
import pytest

from conftest import rgbe_steps, rgbe_truncate

TABLES = [
    # (table, nz, ny, nx, count, shipped key, correlated)
    (0, 1, 128, 128, 262144, "conductor", False),
    (0, 1, 128, 128, 262144, "conductor_correlated", True),
    (1, 128, 64, 128, 131072, "glossy_dielectric", False),
    (2, 128, 16, 256, 65536, "glass", False),
    (3, 128, 16, 256, 65536, "glass_inv", False),
    (4, 96, 32, 32, 65536, "thin_glass", False),
]


def shipped(oracle_mod, key):
    if key == "conductor_correlated":
        path = oracle_mod.LUT_PATH.replace("uncorrelated", "correlated")
        return np.fromfile(path, "<f4").reshape(1, 128, 128)
    a = oracle_mod.load_luts()[key]
    return a.reshape((1,) + a.shape) if a.ndim == 2 else a


def random_texels(rng, nz, ny, nx, n):
    xs, ys, zs = rng.integers(0, nx, n), rng.integers(0, ny, n), rng.integers(0, nz, n)
    # texel 60: wang_hash(61) = 0, a zero xorshift state, every draw 0
    xs[0], ys[0], zs[0] = 60 % nx, (60 // nx) % ny, 60 // (nx * ny)
    return xs, ys, zs


@pytest.mark.parametrize("table,nz,ny,nx,count,key,corr", TABLES, ids=[t[5] for t in TABLES])
def test_oracle_reproduces_shipped_tables(oracle_mod, table, nz, ny, nx, count, key, corr):
    rng = np.random.default_rng(1234 + table + 10 * corr)
    xs, ys, zs = random_texels(rng, nz, ny, nx, 150)
    v = oracle_mod.bake_texels(table, nx, ny, nz, count, xs + ys * nx + zs * nx * ny, correlated=corr)
    ref = shipped(oracle_mod, key)[zs, ny - 1 - ys, xs]
    q = rgbe_truncate(v)
    assert np.isfinite(v).all()
    steps = rgbe_steps(q, ref)
    # grazing texels (cos_theta_o down to 3e-8) are where single samples of
    # huge weight make the last bits of each implementation's arithmetic
    # show; elsewhere the values agree to the file's truncation
    assert (np.abs(steps) <= 1.0).mean() >= 0.98, (key, np.abs(steps).max())
    assert (q == ref).mean() >= 0.97, (key, (q == ref).mean())


def test_flavours_are_told_apart(oracle_mod):
    rng = np.random.default_rng(7)
    xs, ys = rng.integers(0, 128, 150), rng.integers(16, 128, 150)
    corr = shipped(oracle_mod, "conductor_correlated")[0, 127 - ys, xs]
    v = oracle_mod.bake_texels(0, 128, 128, 1, 262144, xs + ys * 128, correlated=False)
    assert (rgbe_truncate(v) == corr).mean() < 0.8


def test_plan(oracle_mod):
    def plan(texels, texels_2d, count):
        f = np.float32
        k = int(np.floor(max(f(1.0), f(1.0e8) / f(texels))))
        n = int(np.ceil(f(count) / f(k)))
        kk = int(np.floor(max(f(1.0), f(1.0e8) / f(texels_2d))))
        return k, n, int(np.ceil(f(count) / f(kk))) * kk
    assert oracle_mod.bake_plan(0, 128, 128, 1, 262144) == plan(128 * 128, 128 * 128, 262144) == (6103, 43, 262429)
    assert oracle_mod.bake_plan(1, 128, 64, 128, 131072) == (95, 1380, 131100)
    assert oracle_mod.bake_plan(2, 256, 16, 128, 65536) == (190, 345, 65550)
    assert oracle_mod.bake_plan(4, 32, 32, 96, 65536) == (1017, 65, 66105)
    # the Fresnel kernel normalises by its 2D count (GGXFresnelDirectionalAlbedo.h:56)
    assert oracle_mod.bake_plan(5, 256, 256, 256, 65536) == plan(256 ** 3, 256 * 256, 65536) == (5, 13108, 65575)


def test_closed_forms(oracle_mod):
    # Fresnel-weighted conductor at F0 = 0 (relative IOR 1: no reflection)
    v = oracle_mod.bake_texels(5, 64, 64, 64, 65536, np.arange(0, 64 * 64, 29))
    assert (v == 0.0).all()
    # perfectly smooth conductor rows (roughness <= 0.01: y = 0, 1 of 128):
    # every sample is (1e9 / cos) / 1e9 * cos
    v = oracle_mod.bake_texels(0, 128, 128, 1, 262144, np.arange(2 * 128))
    assert np.abs(v - 1.0).max() < 2.0e-3
    # glossy dielectric at relative IOR 1: the specular layer is invisible and
    # the MIS mixture integrates the white Lambertian, albedo 1
    xs, ys = np.meshgrid(np.arange(0, 128, 9), np.arange(0, 64, 9))
    v = oracle_mod.bake_texels(1, 128, 64, 128, 131072, (xs + ys * 128).ravel())
    assert np.abs(v - 1.0).max() < 0.01 and abs(v.mean() - 1.0) < 2.0e-3
    # lossless lobes: albedo at most 1 (+ noise) everywhere
    rng = np.random.default_rng(3)
    for table, nz, ny, nx, count in ((0, 1, 128, 128, 16384), (2, 128, 16, 256, 8192), (3, 128, 16, 256, 8192),
                                     (4, 96, 32, 32, 8192)):
        xs, ys, zs = random_texels(rng, nz, ny, nx, 100)
        v = oracle_mod.bake_texels(table, nx, ny, nz, count, xs + ys * nx + zs * nx * ny)
        assert np.isfinite(v).all() and (v > 0).all() and v.max() < 1.02, (table, v.min(), v.max())


def test_lane_partials(oracle_mod):
    """Lanes change only the summation: pairwise partials of the same samples
    (within 0.5 % of the sequential sum, whose drift they do not share), and
    one lane is the sequential sum of the samples in launch order."""
    rng = np.random.default_rng(11)
    tex = rng.integers(0, 128 * 128, 32)
    a = oracle_mod.bake_texels(0, 128, 128, 1, 262144, tex, lanes=1)
    b = oracle_mod.bake_texels(0, 128, 128, 1, 262144, tex, lanes=16)
    assert (np.abs(a - b) / a).max() < 5.0e-3
    c = oracle_mod.bake_texels(0, 128, 128, 1, 262144, tex, lanes=64)
    assert (np.abs(c - b) / b).max() < 5.0e-4  # both round little
    # 2 launches in 2 lanes: p0 + p1 with p_i the launch-i sum alone
    ipk, launches, norm = oracle_mod.bake_plan(0, 128, 128, 1, 2 * 6103)
    assert (ipk, launches) == (6103, 2)
    t = int(tex[0])
    two = oracle_mod.bake_texels(0, 128, 128, 1, 2 * 6103, [t], lanes=2)[0]
    seq = oracle_mod.bake_texels(0, 128, 128, 1, 2 * 6103, [t], lanes=1)[0]
    assert abs(two - seq) / seq < 1.0e-3
