"""GPU parity of RIS direct lighting (sample_lights_RIS, RIS/RIS.h:83-258)
through the C ABI against the oracle: as the direct-light strategy at every
bounce (LSS_RIS_BSDF_AND_LIGHT) and as ReSTIR DI's later-bounce strategy (its
default). The wavefront splits the reservoir: the light candidates stream in
the shading kernel, the BSDF candidate joins after its ray is traced. Bar:
every pixel bit-identical to the oracle, first-hit IDs equal."""
import pytest

from test_gpu_restir import render_both, simplify_materials

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


RIS_VARIANTS = {
    "default": dict(),
    "one_light_no_bsdf": dict(ris_number_of_light_candidates=1, ris_number_of_bsdf_candidates=0),
    "bsdf_only": dict(ris_number_of_light_candidates=0, ris_number_of_bsdf_candidates=1),
    "many_lights": dict(ris_number_of_light_candidates=16),
    "no_candidates": dict(ris_number_of_light_candidates=0, ris_number_of_bsdf_candidates=0),
}


@pytest.mark.parametrize("variant", sorted(RIS_VARIANTS))
def test_ris_strategy_config1(mipt_mod, oracle_mod, torch_mod, variant):
    print(f"\nRIS strategy, config-1 80x80, variant {variant}", flush=True)
    scene = mipt_mod.generate_scene(1, 80, 80)
    s = mipt_mod.default_render_settings(width=80, height=80, nb_bounces=3,
                                         direct_light_sampling_strategy=mipt_mod.MIPT_LSS_RIS_BSDF_AND_LIGHT,
                                         **RIS_VARIANTS[variant])
    fb, _ = render_both(mipt_mod, oracle_mod, scene, s, mipt_mod.default_restir_di_settings(), 4)
    if variant == "no_candidates":
        # no emitter in view and none sampled: black (emission met at later
        # bounces belongs to the direct-light estimator)
        assert fb.max() == 0.0
    else:
        assert fb.mean() > 0.0


def test_ris_strategy_config2_layered(mipt_mod, oracle_mod, torch_mod):
    """Coat, sheen, glass, thin film and diffuse transmission: the candidates'
    BSDF evaluations and the BSDF candidate's sample across every lobe."""
    print("\nRIS strategy, config-2 geometry 96x54, layered materials", flush=True)
    scene = mipt_mod.generate_scene(2, 96, 54)
    s = mipt_mod.default_render_settings(width=96, height=54, nb_bounces=4,
                                         direct_light_sampling_strategy=mipt_mod.MIPT_LSS_RIS_BSDF_AND_LIGHT)
    render_both(mipt_mod, oracle_mod, scene, s, mipt_mod.default_restir_di_settings(), 3)


@pytest.mark.parametrize("later", ["ris", "mis"])
def test_restir_later_bounces(mipt_mod, oracle_mod, torch_mod, later):
    print(f"\nReSTIR DI, later bounces by {later}, config-1 96x96", flush=True)
    scene = mipt_mod.generate_scene(1, 96, 96)
    s = mipt_mod.default_render_settings(width=96, height=96, nb_bounces=3, direct_light_sampling_strategy=5)
    strategy = (mipt_mod.MIPT_RESTIR_DI_LATER_BOUNCES_RIS_BSDF_AND_LIGHT if later == "ris"
                else mipt_mod.MIPT_RESTIR_DI_LATER_BOUNCES_MIS_LIGHT_BSDF)
    rs = mipt_mod.default_restir_di_settings(later_bounces_sampling_strategy=strategy)
    render_both(mipt_mod, oracle_mod, scene, s, rs, 4, check_every_frame=True)


def test_restir_later_bounces_ris_config2(mipt_mod, oracle_mod, torch_mod):
    print("\nReSTIR DI + later-bounce RIS, config-2 geometry 96x54, simplified materials", flush=True)
    scene = simplify_materials(mipt_mod.generate_scene(2, 96, 54))
    s = mipt_mod.default_render_settings(width=96, height=54, nb_bounces=2, direct_light_sampling_strategy=5)
    render_both(mipt_mod, oracle_mod, scene, s, mipt_mod.default_restir_di_settings(), 3)


def test_ris_settings_validation(mipt_mod, torch_mod):
    ctx = mipt_mod.Context(0)
    for bad in (dict(ris_number_of_bsdf_candidates=2), dict(ris_number_of_light_candidates=-1)):
        s = mipt_mod.default_render_settings(width=16, height=16,
                                             direct_light_sampling_strategy=mipt_mod.MIPT_LSS_RIS_BSDF_AND_LIGHT,
                                             **bad)
        with pytest.raises(mipt_mod.MiptError):
            ctx.set_render_settings(s)
    ctx.close()
