"""ReSTIR DI at its stated size: the 1,040,384-triangle scene with 4096
power-sampled emitters at 1920x1080, 32 light + 1 BSDF initial candidates,
temporal reuse (M-cap 20), 2 spatial passes of 5 neighbours in 16 px,
pairwise MIS, 5 bounces, 4 frames (frames 2 and 3 hold pixels whose camera
ray went hit -> miss -> hit under jitter); once more with the fused
spatiotemporal pass fed by the default light presampling (128 subsets of
1024 lights, 8x8 pixel tiles per subset).

The GPU renders the whole frame. Checked:
- over the whole frame, the reservoir invariants (M in [1, M-cap] on the
  pixels with a primary hit; UCW >= 0, -1 after a visibility
  kill, or NaN where the reference's back-facing BSDF candidate weight 0/0
  propagates; the light triangle one of the scene's emitters), a finite
  non-negative framebuffer, and the last frame's first-hit IDs equal to
  the oracle's camera pass on every pixel;
- on two 48x48 tiles (the image centre, and the tile with the most
  hit/miss boundaries), every reservoir buffer and the framebuffer bit for
  bit against the oracle, which renders only the pixels those tiles depend
  on (Renderer.render_frames_in_window).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, H = 1920, 1080
FRAMES = 4
TILE = 48


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32).reshape(a.shape[0], -1)


def _boundary_tile(prims):
    """(x0, x1, y0, y1) of the TILE x TILE tile with the most hit/miss
    neighbours in the last frame's first-hit IDs."""
    hit = (prims.reshape(H, W) >= 0).astype(np.int8)
    edge = np.zeros((H, W), np.int32)
    edge[:, 1:] += hit[:, 1:] != hit[:, :-1]
    edge[1:, :] += hit[1:, :] != hit[:-1, :]
    th, tw = H // TILE, W // TILE
    per = edge[: th * TILE, : tw * TILE].reshape(th, TILE, tw, TILE).sum(axis=(1, 3))
    ty, tx = np.unravel_index(int(np.argmax(per)), per.shape)
    return (tx * TILE, tx * TILE + TILE, ty * TILE, ty * TILE + TILE), int(per[ty, tx])


@pytest.mark.parametrize("case", ["diffuse_metal", "layered", "fused_presampling"])
def test_restir_config2_1080p(mipt_mod, oracle_mod, torch_mod, case):
    materials = "layered" if case == "layered" else "diffuse_metal"
    fused = case == "fused_presampling"
    scene = mipt_mod.generate_scene(2, W, H)
    for m in scene.materials:
        m.specular_transmission = 0.0
        if materials == "diffuse_metal":
            m.coat = m.sheen = m.diffuse_transmission = m.thin_film = m.specular = 0.0
    s = mipt_mod.default_render_settings(width=W, height=H, nb_bounces=5, direct_light_sampling_strategy=5,
                                         direct_light_sampling_base_strategy=1)
    rs = mipt_mod.default_restir_di_settings(number_of_initial_light_candidates=32, m_cap=20,
                                             spatial_number_of_passes=2, do_fused_spatiotemporal=int(fused),
                                             do_light_presampling=int(fused))
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    ctx.set_restir_di_settings(rs)
    ctx.set_render_settings(s)
    ctx.set_camera(scene.camera)
    draws = ctx.draws_per_sample()
    for f in range(FRAMES):
        print(f"restir 1080p {case}: frame {f}", flush=True)
        ctx.render_sample(f)
        torch_mod.cuda.synchronize()
    prims = ctx.read_gbuffer_prims()
    res = {w: ctx.read_reservoirs(w) for w in (0, 1, 2, 3)}
    fb = ctx.read_framebuffer()
    ctx.close()

    hit = prims >= 0
    shaded = res[3]
    assert (shaded["M"][hit] >= 1).all() and (shaded["M"] <= rs.m_cap).all()
    ucw = shaded["ucw"]
    assert ((ucw >= 0) | (ucw == -1) | np.isnan(ucw)).all()
    assert np.isin(shaded["tri"][hit & (ucw > 0)], scene.emissive).all()
    assert (shaded["flags"] & ~0xff == 0).all()
    assert np.isfinite(fb).all() and (fb >= 0).all()

    numbers = lambda f: [mipt_mod.host_random_number(f * draws + k) for k in range(draws)]  # noqa: E731
    cam = oracle_mod.Renderer(scene, s, rs)
    cam.camera_pass(FRAMES - 1, numbers(FRAMES - 1)[0])
    assert np.array_equal(prims, cam.gb_prims), "first-hit IDs differ"
    del cam
    centre = (W // 2 - TILE // 2, W // 2 + TILE // 2, H // 2 - TILE // 2, H // 2 + TILE // 2)
    edge_tile, edges = _boundary_tile(prims)
    assert edges > 0
    for tile in (centre, edge_tile):
        print(f"restir 1080p {case}: oracle on tile {tile}", flush=True)
        R = oracle_mod.Renderer(scene, s, rs)
        R.render_frames_in_window(FRAMES, tile, numbers)
        sel = np.zeros((H, W), bool)
        sel[tile[2]:tile[3], tile[0]:tile[1]] = True
        sel = sel.reshape(-1)
        assert np.array_equal(prims[sel], R.gb_prims[sel]), f"{tile}: first-hit IDs differ"
        for w in (0, 1, 2, 3):
            c = R.read_reservoirs(w)
            same = (_bits(res[w][sel]) == _bits(c[sel])).all(axis=1)
            assert same.all(), (tile, w, np.nonzero(~same)[0][:5])
        g, c = fb.reshape(-1, 3)[sel], R.fb.reshape(-1, 3)[sel]
        assert np.array_equal(g.view(np.uint32), c.view(np.uint32)), f"{tile}: framebuffer differs"
        print(f"tile {tile}: {int(sel.sum())} pixels, {int(hit[sel].sum())} hits, bit-exact over {FRAMES} frames")
