"""The energy-compensation LUT baker on the GPU (mipt_bake_lut) through the
C ABI, against the oracle and against the tables the reference baked and
ships:

- every renderer table (conductor in both masking flavours, glossy
  dielectric, glass entering and exiting, thin glass) baked at the
  reference's settings: 200 random texels (and texel 60, whose xorshift state
  is 0) bit for bit equal to the oracle's restatement, and the whole table,
  through the Radiance file's 8-bit truncation, equal to the shipped one on
  at least 99 % of its texels and within one mantissa step on 99.5 %;
- the Fresnel-weighted table at 64^3: parity, and 0 at F0 = 0;
- lanes per texel 16 (and the library's pick): the same samples summed in
  pairwise partials, bit for bit the oracle's order, over several kernel
  launches that carry the partial sums;
- mipt_set_lut: the five baked tables installed, a configuration-2 frame
  (glass, metals, dielectric specular, coats) renders bit for bit as the
  oracle's with the same tables, and differs from the shipped tables' frame;
  restoring them gives the shipped frame back;
- argument errors.
"""
import numpy as np
import pytest

from conftest import rgbe_steps, rgbe_truncate

pytestmark = pytest.mark.gpu

RENDERER_TABLES = [
    # (name, table, flavour, oracle key)
    ("conductor", 0, 1, "conductor"),
    ("conductor_correlated", 0, 0, None),
    ("glossy_dielectric", 1, 1, "glossy_dielectric"),
    ("glass", 2, 1, "glass"),
    ("glass_inv", 3, 1, "glass_inv"),
    ("thin_glass", 4, 1, "thin_glass"),
]


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


@pytest.fixture(scope="module")
def baked(mipt_mod, torch_mod):
    ctx = mipt_mod.Context(0)
    out = {}
    for name, table, flavour, _ in RENDERER_TABLES:
        s = mipt_mod.default_bake_settings(table, masking_shadowing=flavour)
        print(f"baking {name}", flush=True)
        arr, info = ctx.bake_lut(s)
        print(f"  {name}: {info.ms:.1f} ms, {info.samples / info.ms * 1e-6:.2f} Gsamples/s, "
              f"{info.launches} x {info.iterations_per_launch} samples, {info.kernel_launches} kernel launches",
              flush=True)
        out[name] = (s, arr, info)
    ctx.close()
    return out


def _texels(rng, s, n):
    nz, ny, nx = s.shape
    xs, ys, zs = rng.integers(0, nx, n), rng.integers(0, ny, n), rng.integers(0, nz, n)
    xs[0], ys[0], zs[0] = 60 % nx, (60 // nx) % ny, 60 // (nx * ny)
    return xs, ys, zs


def _oracle_parity(oracle_mod, s, arr, lanes, n, seed):
    nz, ny, nx = s.shape
    xs, ys, zs = _texels(np.random.default_rng(seed), s, n)
    ref = oracle_mod.bake_texels(s.table, nx, ny, nz, s.integration_sample_count, xs + ys * nx + zs * nx * ny,
                                 correlated=s.masking_shadowing == 0, lanes=lanes)
    got = arr[zs, ny - 1 - ys, xs]
    same = got.view(np.uint32) == ref.view(np.uint32)
    assert same.all(), (np.nonzero(~same)[0][:5], got[~same][:5], ref[~same][:5])


@pytest.mark.parametrize("name", [t[0] for t in RENDERER_TABLES])
def test_renderer_tables(mipt_mod, oracle_mod, baked, name):
    s, arr, info = baked[name]
    table, flavour = s.table, s.masking_shadowing
    assert info.lanes_per_texel == 1
    assert (info.iterations_per_launch, info.launches, info.normalizer) == oracle_mod.bake_plan(
        table, s.size_cos_theta, s.size_roughness, s.size_ior, s.integration_sample_count)
    assert np.isfinite(arr).all() and (arr > 0).all()
    _oracle_parity(oracle_mod, s, arr, 1, 200, 100 + table)
    shipped = mipt_mod.shipped_lut(table, flavour)
    steps = rgbe_steps(rgbe_truncate(arr), shipped)
    eq, within = (steps == 0).mean(), (np.abs(steps) <= 1).mean()
    print(f"\n{name}: {eq * 100:.3f} % of {arr.size} texels equal the shipped table, "
          f"{within * 100:.3f} % within one step, max {np.abs(steps).max():.0f} steps")
    assert eq >= 0.99 and within >= 0.995


def test_fresnel_table(mipt_mod, oracle_mod, torch_mod):
    s = mipt_mod.default_bake_settings(mipt_mod.MIPT_LUT_GGX_FRESNEL, size_cos_theta=64, size_roughness=64,
                                       size_ior=64)
    ctx = mipt_mod.Context(0)
    arr, info = ctx.bake_lut(s)
    ctx.close()
    assert (info.iterations_per_launch, info.launches, info.normalizer) == oracle_mod.bake_plan(5, 64, 64, 64, 65536)
    assert (arr[0] == 0.0).all() and (arr[1:] > 0).all() and arr.max() <= 1.0
    _oracle_parity(oracle_mod, s, arr, 1, 200, 7)


@pytest.mark.parametrize("lanes", [16, 0])
def test_lane_partials(mipt_mod, oracle_mod, baked, lanes):
    base_s, base, _ = baked["conductor"]
    s = mipt_mod.default_bake_settings(mipt_mod.MIPT_LUT_GGX_CONDUCTOR, lanes_per_texel=lanes)
    ctx = mipt_mod.Context(0)
    arr, info = ctx.bake_lut(s)
    ctx.close()
    print(f"\nconductor with {info.lanes_per_texel} lanes per texel: {info.ms:.1f} ms, "
          f"{info.kernel_launches} kernel launches")
    assert info.lanes_per_texel > 1 and info.kernel_launches > 1
    _oracle_parity(oracle_mod, s, arr, info.lanes_per_texel, 200, 31)
    # the same samples, summed with less rounding drift than one lane's
    assert (np.abs(arr - base) / base).max() < 5.0e-3


def _render(mipt_mod, scene, s, tables=None, restore=False):
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    ctx.set_render_settings(s)
    ctx.set_camera(scene.camera)
    if tables:
        for t, arr in tables.items():
            ctx.set_lut(t, arr)
    ctx.render_sample(0)
    if restore:
        for t in tables:
            ctx.set_lut(t, None)
        ctx.render_sample(0)
    fb = ctx.read_framebuffer()
    ctx.close()
    return fb


def test_installed_tables_render(mipt_mod, oracle_mod, baked, torch_mod):
    scene = mipt_mod.generate_scene(2, 160, 90)
    s = mipt_mod.default_render_settings(width=160, height=90, nb_bounces=4)
    names = ["conductor", "glossy_dielectric", "glass", "glass_inv", "thin_glass"]
    tables = {baked[n][0].table: baked[n][1] for n in names}
    fb = _render(mipt_mod, scene, s, tables)
    lut = {n: baked[n][1].reshape(oracle_mod.load_luts()[n].shape) for n in names}
    R = oracle_mod.Renderer(scene, s, lut=lut)
    R.render_sample(0, mipt_mod.host_random_number(0), mipt_mod.host_random_number(1))
    assert np.array_equal(fb.view(np.uint32), R.fb.view(np.uint32))
    shipped_fb = _render(mipt_mod, scene, s)
    assert not np.array_equal(fb, shipped_fb)
    restored = _render(mipt_mod, scene, s, tables, restore=True)
    assert np.array_equal(restored.view(np.uint32), shipped_fb.view(np.uint32))


def test_errors(mipt_mod, torch_mod):
    ctx = mipt_mod.Context(0)
    bad = [
        dict(table=-1),
        dict(size_ior=2),  # conductor is 2D
        dict(size_cos_theta=1),
        dict(integration_sample_count=0),
        dict(masking_shadowing=2),
        dict(lanes_per_texel=3),
        dict(lanes_per_texel=128),
    ]
    for kw in bad:
        s = mipt_mod.default_bake_settings(0)
        for k, v in kw.items():
            setattr(s, k, v)
        with pytest.raises(mipt_mod.MiptError):
            ctx.bake_lut(s)
    with pytest.raises(mipt_mod.MiptError):
        ctx.set_lut(mipt_mod.MIPT_LUT_GGX_CONDUCTOR, np.zeros((64, 128), np.float32))
    with pytest.raises(mipt_mod.MiptError):
        ctx.set_lut(mipt_mod.MIPT_LUT_GGX_FRESNEL, np.zeros((4, 4, 4), np.float32))
    ctx.close()
