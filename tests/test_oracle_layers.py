"""Pins of the oracle's layered Principled lobes (dielectric specular over the
diffuse base, coat, sheen) by properties the mathematics fixes:

- white furnace: with a white base every layer passes on what it does not
  reflect (1 - F, 1 - sheen reflectance), so the compensated glossy base, the
  coat over a white base and the sheen over a white base all have directional
  albedo 1 (up to the LUT's texel quantisation). A wrong LUT axis, texel or
  layer attenuation moves the albedo off 1;
- the LTC sheen distribution integrates to 1 over the hemisphere, and its
  reflectance is the tabulated Ri;
- sample() draws directions with the density pdf() reports (histogram test),
  returns the eval at the sampled direction, and eval's pdf equals pdf().
"""
import numpy as np
import pytest

from test_oracle_shading import _albedo_is, _albedo_mc, _frame


def _dielectric(mipt_mod, roughness, ior, specular=1.0, base=1.0):
    m = mipt_mod.default_material()
    for k in range(3):
        m.base_color[k] = base
    m.metallic = 0.0
    m.specular = specular
    m.roughness = roughness
    m.ior = ior
    return m


@pytest.mark.parametrize("roughness,ior,theta", [(0.2, 1.5, 0.3), (0.5, 1.45, 0.9), (0.9, 1.8, 0.6),
                                                 (0.35, 1.25, 1.2)])
def test_glossy_base_white_furnace(mipt_mod, oracle_mod, roughness, ior, theta):
    bsdf = oracle_mod.Bsdf(_dielectric(mipt_mod, roughness, ior))
    view, n = _frame(theta)
    albedo, _ = _albedo_is(bsdf, view, n, 12000, state=3)
    assert np.allclose(albedo, 1.0, atol=0.04), albedo
    # without the specular energy compensation the glossy base loses energy
    m = _dielectric(mipt_mod, roughness, ior)
    m.do_specular_energy_compensation = 0
    plain, _ = _albedo_is(oracle_mod.Bsdf(m), view, n, 12000, state=3)
    assert plain[0] < albedo[0]


@pytest.mark.parametrize("coat_roughness,coat_ior,darkening", [(0.1, 1.5, 1.0), (0.5, 1.6, 0.0), (0.8, 1.3, 1.0)])
def test_coat_white_furnace(mipt_mod, oracle_mod, coat_roughness, coat_ior, darkening):
    m = _dielectric(mipt_mod, 0.5, 1.5, specular=0.0)
    m.coat = 1.0
    m.coat_roughness = coat_roughness
    m.coat_ior = coat_ior
    m.coat_darkening = darkening
    view, n = _frame(0.5)
    albedo, _ = _albedo_is(oracle_mod.Bsdf(m), view, n, 12000, state=5)
    assert np.allclose(albedo, 1.0, atol=0.04), albedo


def test_coat_absorption_darkens(mipt_mod, oracle_mod):
    """An absorbing coat medium transmits exp(-(1 - sqrt(a)^d) t) of the base."""
    m = _dielectric(mipt_mod, 0.5, 1.5, specular=0.0)
    m.coat = 1.0
    m.coat_roughness = 0.3
    view, n = _frame(0.2)
    clear, _ = _albedo_is(oracle_mod.Bsdf(m), view, n, 6000, state=9)
    for k, a in enumerate((0.5, 0.8, 1.0)):
        m.coat_medium_absorption[k] = a
    tinted, _ = _albedo_is(oracle_mod.Bsdf(m), view, n, 6000, state=9)
    assert tinted[0] < tinted[1] < tinted[2] <= clear[2] + 1e-6


@pytest.mark.parametrize("sheen_roughness", [0.1, 0.5, 1.0])
def test_sheen_white_furnace_and_ltc_normalisation(mipt_mod, oracle_mod, sheen_roughness):
    m = _dielectric(mipt_mod, 0.5, 1.5, specular=0.0)
    m.sheen = 1.0
    m.sheen_roughness = sheen_roughness
    bsdf = oracle_mod.Bsdf(m)
    view, n = _frame(0.6)
    albedo, _ = _albedo_is(bsdf, view, n, 12000, state=11)
    assert np.allclose(albedo, 1.0, atol=0.03), albedo
    # sheen and diffuse are each sampled with probability 1/2: the pdf over the
    # hemisphere integrates to 1/2 (1 + integral of the LTC) = 1
    _, pdf_int = _albedo_mc(bsdf, view, n, 8000, seed=2)
    assert abs(pdf_int - 1.0) < 0.04
    # a black base isolates the sheen lobe: its albedo is the LTC reflectance Ri
    # (sheen_ltc_32x32.f32, nearest texel of (cos_theta, roughness))
    for k in range(3):
        m.base_color[k] = 0.0
    only, _ = _albedo_mc(oracle_mod.Bsdf(m), view, n, 8000, seed=3)
    tab = oracle_mod.load_luts()["sheen_ltc"].reshape(-1, 3)
    cos_t = np.float32(view[2])
    r = np.float32(int(sheen_roughness * 255.0) / 255.0)
    x, y = int(cos_t * 31.0), int((np.float32(1.0) - r) * 31.0)
    assert abs(only[0] - tab[x + 32 * y, 2]) < 0.03


def _full_material(mipt_mod, rng):
    m = mipt_mod.default_material()
    for k in range(3):
        m.base_color[k] = rng.uniform(0.1, 0.9)
        m.specular_color[k] = rng.uniform(0.5, 1.0)
        m.sheen_color[k] = rng.uniform(0.5, 1.0)
    m.roughness = rng.uniform(0.15, 1.0)
    m.metallic = rng.uniform(0, 1)
    m.specular = rng.uniform(0, 1)
    m.specular_tint = rng.uniform(0, 1)
    m.coat = rng.uniform(0, 1)
    m.coat_roughness = rng.uniform(0.1, 1.0)
    m.sheen = rng.uniform(0, 1)
    m.sheen_roughness = rng.uniform(0.05, 1.0)
    m.anisotropy = rng.uniform(0, 1)
    m.anisotropy_rotation = rng.uniform(0, 1)
    m.ior = rng.uniform(1.1, 2.0)
    return m


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_layered_sampling_matches_pdf(mipt_mod, oracle_mod, seed):
    rng = np.random.default_rng(seed)
    bsdf = oracle_mod.Bsdf(_full_material(mipt_mod, rng))
    view, n = _frame(rng.uniform(0.1, 1.2))
    state = 1000 + seed
    count = 20000
    cos_s, phi_s = [], []
    for _ in range(count):
        d, f, pdf, state = bsdf.sample(view, n, n, state)
        if pdf > 0 and d[2] > 0:
            f2, pdf2 = bsdf.eval(view, n, n, d)
            assert abs(pdf2 - pdf) <= 1e-5 * max(1.0, pdf)
            cos_s.append(d[2])
            phi_s.append(np.arctan2(d[1], d[0]))
    # histogram over (cos theta, phi) against the pdf, by uniform hemisphere MC
    ce, pe = np.linspace(0, 1, 6), np.linspace(-np.pi, np.pi, 7)
    hist = np.histogram2d(cos_s, phi_s, [ce, pe])[0] / count
    N = 30000
    u1, u2 = rng.random(N), rng.random(N)
    exp = np.zeros((5, 6))
    for z, ph in zip(u1, 2 * np.pi * u2 - np.pi):
        s = np.sqrt(1 - z * z)
        d = np.array([s * np.cos(ph), s * np.sin(ph), z], np.float32)
        _, pdf = bsdf.eval(view, n, n, d)
        exp[min(int(z * 5), 4), min(int((ph + np.pi) / (2 * np.pi) * 6), 5)] += pdf * 2 * np.pi / N
    assert np.abs(hist - exp).max() < 0.02, np.abs(hist - exp).max()


def test_eval_pdf_equals_pdf_function(mipt_mod, oracle_mod):
    rng = np.random.default_rng(7)
    for _ in range(40):
        bsdf = oracle_mod.Bsdf(_full_material(mipt_mod, rng))
        view, n = _frame(rng.uniform(0.0, 1.4))
        for _ in range(25):
            d = rng.normal(size=3)
            d[2] = abs(d[2])
            d = (d / np.linalg.norm(d)).astype(np.float32)
            _, pdf = bsdf.eval(view, n, n, d)
            assert abs(bsdf.pdf(view, n, n, d) - pdf) <= 1e-6 * max(1.0, pdf)
