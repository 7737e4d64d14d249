"""GPU parity of the glass and diffuse-transmission lobes inside the
path-tracing and ReSTIR DI passes: the per-path volume state (interior stack
pushed at every hit, popped by reflections and refractions out, distance in
the medium, sampled wavelength), priority-skipped boundaries re-traced in
extra rounds, Beer-Lambert absorption and the dispersion filter. Every frame
is compared with the CPU oracle bit for bit, like the opaque scenes
(test_gpu_pathtrace.py, test_gpu_restir.py).
"""
import numpy as np
import pytest

from test_gpu_pathtrace import _check, _render_gpu
from test_gpu_restir import render_both

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _set_glass(m, roughness=0.0, ior=1.5, priority=0, absorption=None, at_distance=1.0, thin=0, dt=0.0,
               dispersion=0.0, abbe=20.0, coat=0.0, base=(1.0, 1.0, 1.0)):
    for k in range(3):
        m.base_color[k] = base[k]
    m.metallic = 0.0
    m.specular = 0.5
    m.roughness = roughness
    m.ior = ior
    m.specular_transmission = 1.0
    m.diffuse_transmission = dt
    m.dielectric_priority = priority
    m.thin_walled = thin
    m.absorption_at_distance = at_distance
    m.dispersion_scale = dispersion
    m.dispersion_abbe_number = abbe
    m.coat = coat
    if absorption is not None:
        for k in range(3):
            m.absorption_color[k] = absorption[k]


def nested_glass_scene(mipt_mod, w, h, dispersive=True):
    """Config-1 geometry (Cornell box, two ceiling lights, 20 spheres, each
    sphere its own material) with six spheres turned to glass and three of them
    holding a concentric inner sphere (scaled 0.55 about the sphere's centroid):
    a lower-priority inner medium (skipped while inside the outer one), an
    equal-priority one (entered from the outer medium) and a higher-priority
    one (it wins over the outer medium)."""
    scene = mipt_mod.generate_scene(1, w, h)
    mats = list(scene.materials)
    mi = scene.material_indices
    sphere = [k for k in range(len(mats)) if (mi == k).sum() == 5120]
    g = sphere[:6]
    _set_glass(mats[g[0]], 0.0, 1.5, priority=2)                                        # smooth, holds a skipped core
    _set_glass(mats[g[1]], 0.3, 1.45, absorption=(0.5, 0.75, 0.9), at_distance=0.6)     # rough, absorbing
    _set_glass(mats[g[2]], 0.05, 1.6, thin=1, base=(0.9, 0.95, 1.0))                    # thin-walled
    _set_glass(mats[g[3]], 0.4, 1.3, dt=0.6, base=(0.8, 0.6, 0.5))                      # glass + diffuse transmission
    if dispersive:
        _set_glass(mats[g[4]], 0.02, 1.55, dispersion=0.8, abbe=25.0, priority=1)       # dispersive, holds a core
    else:
        _set_glass(mats[g[4]], 0.02, 1.55, priority=1)
    _set_glass(mats[g[5]], 0.2, 1.5, coat=1.0, priority=1)                              # coated glass, holds a core
    verts = scene.vertices
    idx = scene.indices
    new_v, new_i, new_m = [verts], [idx], [mi]
    inner = []
    for outer, (ior, prio, absorb) in zip((g[0], g[4], g[5]), ((1.33, 1, None), (1.8, 1, (0.6, 0.6, 0.9)),
                                                               (1.2, 3, None))):
        tris = idx[mi == outer]
        used = np.unique(tris)
        c = verts[used].astype(np.float64).mean(axis=0)
        remap = -np.ones(verts.shape[0], np.int64)
        base = sum(v.shape[0] for v in new_v)
        remap[used] = base + np.arange(used.size)
        new_v.append((c + 0.55 * (verts[used].astype(np.float64) - c)).astype(np.float32))
        new_i.append(remap[tris].astype(np.int32))
        m = mipt_mod.default_material()
        _set_glass(m, 0.1, ior, priority=prio, absorption=absorb, at_distance=0.4)
        mats.append(m)
        new_m.append(np.full(tris.shape[0], len(mats) - 1, np.int32))
        inner.append(len(mats) - 1)
    table = (type(mats[0]) * len(mats))(*mats)
    return mipt_mod.Scene(np.concatenate(new_v), np.concatenate(new_i), np.concatenate(new_m), table, scene.emissive,
                          scene.camera, w, h)


def test_config1_nested_glass(mipt_mod, oracle_mod, torch_mod):
    """Smooth, rough, absorbing, thin-walled, dispersive, coated glass and
    diffuse transmission, with nested media of lower, equal and higher
    priority; 8 bounces so that paths cross several interfaces."""
    scene = nested_glass_scene(mipt_mod, 96, 96)
    s = mipt_mod.default_render_settings(width=96, height=96, nb_bounces=8)
    fb, cfb = _check(mipt_mod, oracle_mod, scene, s, 3)


def test_config1_glass_without_roulette(mipt_mod, oracle_mod, torch_mod):
    """The same scene with Russian roulette off and regularisation off: long
    paths through the glass, the delta (smooth) lobes unregularised."""
    scene = nested_glass_scene(mipt_mod, 64, 64)
    s = mipt_mod.default_render_settings(width=64, height=64, nb_bounces=10, do_russian_roulette=0,
                                         do_microfacet_regularization=0)
    _check(mipt_mod, oracle_mod, scene, s, 2)


def test_config2_full_materials(mipt_mod, oracle_mod, torch_mod):
    """The 1M-triangle scene with every principled parameter drawn per object,
    specular transmission included; uniform NEE + MIS, 5 bounces."""
    scene = mipt_mod.generate_scene(2, 160, 90)
    s = mipt_mod.default_render_settings(width=160, height=90, nb_bounces=5)
    _check(mipt_mod, oracle_mod, scene, s, 2)


def test_restir_config1_glass(mipt_mod, oracle_mod, torch_mod):
    """ReSTIR DI with glass surfaces at the first hit (non-dispersive)."""
    scene = nested_glass_scene(mipt_mod, 80, 80, dispersive=False)
    s = mipt_mod.default_render_settings(width=80, height=80, nb_bounces=6, direct_light_sampling_strategy=5)
    rs = mipt_mod.default_restir_di_settings()
    fb, R = render_both(mipt_mod, oracle_mod, scene, s, rs, 4)
    assert fb.mean() > 0.0


def test_restir_with_dispersion_is_refused(mipt_mod, torch_mod):
    scene = nested_glass_scene(mipt_mod, 32, 32)
    s = mipt_mod.default_render_settings(width=32, height=32, direct_light_sampling_strategy=5)
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    ctx.set_render_settings(s)
    ctx.set_camera(scene.camera)
    with pytest.raises(mipt_mod.MiptError) as e:
        ctx.render_sample(0)
    assert "dispersive" in str(e.value)
    ctx.close()


def test_config1_thin_film(mipt_mod, oracle_mod, torch_mod):
    """Thin-film interference on a metal, a dielectric specular layer, a
    coated base and a glass sphere (hue shift, IOR override with kappa, partial
    film weight), 6 bounces."""
    scene = nested_glass_scene(mipt_mod, 96, 96, dispersive=False)
    mats = scene.materials
    mi = scene.material_indices
    sphere = [k for k in range(len(mats)) if (mi == k).sum() == 5120]
    films = [(sphere[6], dict(metallic=1.0, roughness=0.3, thin_film=1.0, thin_film_thickness=420.0,
                              thin_film_ior=1.4, thin_film_hue_shift_degrees=40.0)),
             (sphere[7], dict(metallic=0.0, specular=1.0, roughness=0.2, thin_film=0.7, thin_film_thickness=250.0,
                              thin_film_ior=1.6)),
             (sphere[8], dict(metallic=0.0, specular=0.8, coat=1.0, thin_film=1.0, thin_film_thickness=600.0,
                              thin_film_ior=1.3, thin_film_do_ior_override=1, thin_film_base_ior_override=0.2,
                              thin_film_kappa_3=3.0)),
             (sphere[2], dict(thin_film=0.5, thin_film_thickness=500.0, thin_film_ior=1.35))]  # thin glass
    for k, params in films:
        for a, v in params.items():
            setattr(mats[k], a, v)
    s = mipt_mod.default_render_settings(width=96, height=96, nb_bounces=6)
    _check(mipt_mod, oracle_mod, scene, s, 3)


def test_glass_changes_the_image(mipt_mod, torch_mod):
    """The glass spheres transmit: the nested-glass scene differs from the same
    geometry with the glass made opaque, and both stay finite."""
    scene = nested_glass_scene(mipt_mod, 64, 64)
    s = mipt_mod.default_render_settings(width=64, height=64, nb_bounces=6)
    fb, _, _, rej = _render_gpu(mipt_mod, scene, s, 4)
    for m in scene.materials:
        m.specular_transmission = 0.0
        m.diffuse_transmission = 0.0
    fb2, _, _, _ = _render_gpu(mipt_mod, scene, s, 4)
    assert np.isfinite(fb).all() and np.isfinite(fb2).all()
    assert np.abs(fb - fb2).mean() > 1e-3 * fb2.mean()
