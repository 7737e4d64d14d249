"""Pins of the oracle's thin-film interference Fresnel (Belcour & Barla
2017, ThinFilm.h) by what the model fixes:

- a film whose IOR matches the ambient medium has no first interface: the
  Airy sum reduces to the bare Fresnel reflectance of the base under it
  (float64 textbook formula, to 5e-4 in binary32; the XYZ -> sRGB rows sum
  to 1 within 3e-5),
  for a dielectric base and for a conductor base (complex IOR override);
- a 120 degree hue rotation about the grey axis permutes the channels;
- the result lies in [0, 1], is grey-free (iridescent) for a film of a few
  hundred nm, and is 0 without a film;
- with thin_film = 1 the specular layer's Fresnel is the thin-film one (the
  BSDF evaluated at normal incidence equals it times the specular lobe's
  value for F = 1).
"""
import math

import numpy as np
import pytest


def _fresnel_dielectric(cos_i, eta):
    sin_t2 = (1.0 - cos_i * cos_i) / (eta * eta)
    if sin_t2 >= 1.0:
        return 1.0
    cos_t = math.sqrt(1.0 - sin_t2)
    rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t)
    rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t)
    return 0.5 * (rs * rs + rp * rp)


def _fresnel_conductor(cos_i, n, k):
    """Unpolarised reflectance at a conductor of complex index n + ik (float64)."""
    eta = complex(n, k)
    sin2 = 1.0 - cos_i * cos_i
    cos_t = np.sqrt(1.0 - sin2 / (eta * eta))
    rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t)
    rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t)
    return 0.5 * (abs(rs) ** 2 + abs(rp) ** 2)


def _film(mipt_mod, film_ior=1.33, thickness=400.0, base_ior=1.5, override=None, hue=0.0):
    m = mipt_mod.default_material()
    m.thin_film = 1.0
    m.thin_film_ior = film_ior
    m.thin_film_thickness = thickness
    m.ior = base_ior
    m.thin_film_hue_shift_degrees = hue
    if override is not None:
        m.thin_film_do_ior_override = 1
        m.thin_film_base_ior_override, m.thin_film_kappa_3 = override
    return m


ROW_SUMS = np.array([2.3646381 - 0.8965361 - 0.4680737, -0.5151664 + 1.4264 + 0.0887608,
                     0.0052037 - 0.0144081 + 1.0092106])


@pytest.mark.parametrize("cos_i", [1.0, 0.8, 0.5, 0.2])
def test_index_matched_film_is_the_bare_dielectric(mipt_mod, oracle_mod, cos_i):
    m = _film(mipt_mod, film_ior=1.3, base_ior=1.7)
    F = oracle_mod.thin_film_fresnel(m, 1.3, cos_i).astype(np.float64)
    expect = ROW_SUMS * _fresnel_dielectric(cos_i, 1.7 / 1.3)
    assert np.allclose(F, expect, rtol=5e-4), (F, expect)  # binary32 cancellation at grazing angles


@pytest.mark.parametrize("cos_i", [1.0, 0.6, 0.3])
def test_index_matched_film_over_a_conductor(mipt_mod, oracle_mod, cos_i):
    # silver-like base (n = 0.15, k = 3.6) under a film matched to the ambient
    m = _film(mipt_mod, film_ior=1.2, override=(0.15, 3.6))
    F = oracle_mod.thin_film_fresnel(m, 1.2, cos_i).astype(np.float64)
    # the model takes the base's kappa unscaled by the film IOR (ThinFilm.h:149)
    expect = ROW_SUMS * _fresnel_conductor(cos_i, 0.15 / 1.2, 3.6)
    assert np.allclose(F, expect, rtol=1e-4), (F, expect)


def test_hue_rotation_by_120_degrees_permutes_channels(oracle_mod):
    c = np.array([0.2, 0.5, 0.9], np.float32)
    assert np.allclose(oracle_mod.hue_shift(c, 120.0), [0.9, 0.2, 0.5], atol=1e-6)
    assert np.allclose(oracle_mod.hue_shift(c, 240.0), [0.5, 0.9, 0.2], atol=1e-6)
    assert np.array_equal(oracle_mod.hue_shift(c, 0.0), c)


def test_iridescence_range_and_colour(mipt_mod, oracle_mod):
    cols = []
    for t in np.linspace(100.0, 900.0, 17):
        for cos_i in (0.95, 0.5, 0.1):
            F = oracle_mod.thin_film_fresnel(_film(mipt_mod, thickness=float(t)), 1.0, cos_i)
            assert ((F >= 0.0) & (F <= 1.0)).all()
            cols.append(F)
    cols = np.array(cols)
    assert (cols.max(axis=1) - cols.min(axis=1)).max() > 0.01  # coloured, not grey
    m = _film(mipt_mod)
    m.thin_film = 0.0
    assert np.array_equal(oracle_mod.thin_film_fresnel(m, 1.0, 0.7), np.zeros(3, np.float32))


def test_full_film_replaces_the_specular_fresnel(mipt_mod, oracle_mod):
    """A dielectric specular layer (no diffuse, white specular colour) at
    normal incidence: BSDF(thin film) / BSDF(no film) = F_film / F_dielectric
    per channel, compensation off."""
    def mat(film):
        m = _film(mipt_mod, thickness=350.0, base_ior=1.5)
        m.thin_film = 1.0 if film else 0.0
        for k in range(3):
            m.base_color[k] = 0.0
        m.metallic = 0.0
        m.specular = 1.0
        m.roughness = 0.4
        m.do_specular_energy_compensation = 0
        return m

    n = np.array([0.0, 0.0, 1.0], np.float32)
    v = np.array([0.0, 0.0, 1.0], np.float32)
    f_film, _ = oracle_mod.Bsdf(mat(True)).eval(v, n, n, v)
    f_plain, _ = oracle_mod.Bsdf(mat(False)).eval(v, n, n, v)
    F_film = oracle_mod.thin_film_fresnel(mat(True), 1.0, 1.0)
    F_plain = _fresnel_dielectric(1.0, 1.5)
    assert np.allclose(f_film / f_plain, F_film / F_plain, rtol=1e-4), (f_film / f_plain, F_film / F_plain)
