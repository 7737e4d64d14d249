"""Pins the oracle's RIS direct lighting (sample_lights_RIS,
RIS/RIS.h:83-258) by what the mathematics fixes: resampled importance
sampling with balance-heuristic candidate weights and the unbiased contribution
weight wsum / target is an unbiased estimator of the same direct lighting as
NEE + MIS, for any number of light and BSDF candidates. A candidate weight
without its MIS term, with the candidate counts swapped, or a UCW without its
1 / target shows up as a bias of several % against the MIS mean; a dropped
shadow test on the kept light sample shows up as a brighter image.

The same holds for ReSTIR DI's later bounces (RIS by default, MIS light +
BSDF as the alternative): both estimate the same indirect-bounce direct light."""
import numpy as np
import pytest

W = H = 48


def _render(mipt_mod, oracle_mod, scene, frames, strategy, bounces=0, restir=None, **settings):
    s = mipt_mod.default_render_settings(width=W, height=H, nb_bounces=bounces,
                                         direct_light_sampling_strategy=strategy, **settings)
    rs = mipt_mod.default_restir_di_settings(**(restir or {}))
    R = oracle_mod.Renderer(scene, s, rs)
    draws = 2 if strategy != 5 else 3 + 2 * rs.do_temporal_reuse_pass + \
        rs.spatial_number_of_passes * rs.do_spatial_reuse_pass
    for f in range(frames):
        R.render_frame(f, [mipt_mod.host_random_number(f * draws + k) for k in range(draws)])
    return R.fb64.reshape(-1, 3) / frames, R


@pytest.fixture(scope="module")
def scene48(mipt_mod):
    return mipt_mod.generate_scene(1, W, H)


@pytest.fixture(scope="module")
def mis_mean(mipt_mod, oracle_mod, scene48):
    img, _ = _render(mipt_mod, oracle_mod, scene48, 512, 3)
    return img.mean(axis=0)


@pytest.mark.parametrize("lights,bsdfs", [(4, 1), (1, 0), (0, 1), (8, 1), (2, 0)])
def test_ris_is_unbiased(mipt_mod, oracle_mod, scene48, mis_mean, lights, bsdfs):
    frames = 1024 if lights == 0 else 384
    img, R = _render(mipt_mod, oracle_mod, scene48, frames, 4, ris_number_of_light_candidates=lights,
                     ris_number_of_bsdf_candidates=bsdfs)
    assert R.rejected == 0
    assert np.isfinite(img).all() and (img >= 0).all()
    tol = 0.03 if lights == 0 else 0.01
    assert np.allclose(img.mean(axis=0), mis_mean, rtol=tol), (img.mean(axis=0), mis_mean)


def test_ris_differs_from_mis_per_sample(mipt_mod, oracle_mod, scene48):
    """RIS is a different estimator: the same random numbers give another image."""
    a, _ = _render(mipt_mod, oracle_mod, scene48, 1, 3)
    b, _ = _render(mipt_mod, oracle_mod, scene48, 1, 4)
    assert not np.array_equal(a, b)


def test_ris_without_candidates_is_black(mipt_mod, oracle_mod, scene48):
    """No candidates: an empty reservoir (UCW 0) at every bounce, and the
    emission that later bounces meet belongs to the direct-light estimator, so
    with no emitter in view the image is black."""
    img, R = _render(mipt_mod, oracle_mod, scene48, 2, 4, bounces=2, ris_number_of_light_candidates=0,
                     ris_number_of_bsdf_candidates=0)
    assert not np.isin(R.gb_prims, scene48.emissive).any()
    assert (img == 0).all()


def test_restir_later_bounces_ris_and_mis_agree(mipt_mod, oracle_mod, scene48):
    """ReSTIR DI at bounce 0 with RIS or MIS at the later bounces: the same
    one-bounce image in the mean."""
    common = dict(do_temporal_reuse_pass=0, do_spatial_reuse_pass=0)
    ris, _ = _render(mipt_mod, oracle_mod, scene48, 256, 5, bounces=1, restir=common)
    mis, _ = _render(mipt_mod, oracle_mod, scene48, 256, 5, bounces=1,
                     restir=dict(common, later_bounces_sampling_strategy=2))
    zero, _ = _render(mipt_mod, oracle_mod, scene48, 256, 5, bounces=0, restir=common)
    assert np.allclose(ris.mean(axis=0), mis.mean(axis=0), rtol=0.01)
    # the bounce adds light (the later-bounce estimator is active)
    assert (ris.mean(axis=0) > zero.mean(axis=0) * 1.05).all()
